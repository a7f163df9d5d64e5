"""Screening: queries against a conformer library resident in HBM, with continuous batching.

Front-end restating the reference's screening contract (shepherd_score/screen.py:74, 1038-1192):
``screen(query, library, mode, top_k, **align_kwargs) -> [Hit(score, id, transform)]`` sorted by
score, descending.  The reference streams ``ProfileStore`` shards from disk through per-query
aligner calls; here the library is packed once into HBM and served by the native screening
service of libfss_hip.so (``fss_lib_*`` / ``fss_screen_*`` in include/fss.h), which this module
wraps:

* admission by memory: a query takes one of ``max_active`` top-k slots and, for its own rows,
  ranges of four device row pools (shape rows, pharmacophores, ShaEP surface, ShaEP atoms)
  reserved when the service is created and sized from the HBM budget
  (:func:`plan_query_pools`); queries wait in FIFO order until a slot and room for their rows
  are free, and no later query overtakes a waiting one (:func:`admit_fifo`);
* continuous batching: every launch is filled from the admitted queries' remaining library
  ranges, oldest first (:func:`plan_launch` / ``fss_plan_launch``): the tail of one query and
  the head of the next share a launch, and a query that arrives while others are on the GPU
  joins as soon as capacity frees;
* per-pair results do not depend on the co-scheduled pairs (fixed tiling, per-pair seeds and
  early stop, fixed-order reductions) and the device-side top-k merge orders by (score
  descending, library index ascending), so a query's hits are identical alone and in the mix.
"""
from __future__ import annotations

import collections
import ctypes
import os
import threading
from concurrent.futures import Future
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .align import OptimSpec, ScoreSpec
from .constants import LAM_SCALING
from .packing import CloudSet, PharmSet, ShaepSet
from .seeds import quats_from_fibo

Hit = collections.namedtuple("Hit", ["score", "id", "transform"])

_KIND = {"vol": _lib.MODE_SHAPE, "surf": _lib.MODE_SHAPE, "vol_esp": _lib.MODE_SHAPE_ESP,
         "surf_esp": _lib.MODE_SHAPE_ESP, "pharm": _lib.MODE_PHARM, "vol_color": _lib.MODE_SHAPE_PHARM,
         "vol_and_surf_esp": _lib.MODE_ESP_COMBO}


def restart_table(B: int) -> np.ndarray:
    """Screening restarts: identity, then rotations onto B-1 Fibonacci-sphere points."""
    q = np.zeros((B, 4), np.float32)
    q[0, 0] = 1.0
    if B > 1:
        q[1:] = quats_from_fibo(B - 1).astype(np.float32)
    return q


class ConformerLibrary:
    """Packed conformers on one device: shape channel (atoms or surface points [+ ESP]),
    pharmacophores, and either the ShaEP channel (surface points with ESP, atoms with hydrogens,
    partial charges and radii) or surface rows alone (points with ESP), with integer ids.  The
    surf / surf_esp screens read a library's surface rows when it has them, so one library with
    heavy atoms, pharmacophores and surfaces serves the atom and the surface modes."""

    def __init__(self, shape: Optional[CloudSet] = None, pharm: Optional[PharmSet] = None,
                 ids: Optional[np.ndarray] = None, shaep: Optional[ShaepSet] = None,
                 surface: Optional[CloudSet] = None):
        if shaep is not None and surface is not None:
            raise ValueError("the ShaEP channel holds the surface rows: pass shaep or surface, not both")
        self.shape, self.pharm, self.shaep, self.surface = shape, pharm, shaep, surface
        some = [c for c in (shape, pharm, shaep, surface) if c is not None]
        if not some:
            raise ValueError("a library needs at least one channel")
        n = some[0].n
        if any(c.n != n for c in some):
            raise ValueError("every channel must hold the same conformers")
        self.n = n
        self.ids = np.arange(n, dtype=np.int64) if ids is None else np.asarray(ids)
        self._native = None
        self._native_bytes = None
        # surface points per conformer when every surface cloud has the same count (the
        # reference's store property that picks ALPHA for surf / surf_esp), and whether the
        # conformers were centred on their heavy-atom mean when stored
        self.num_surf_points = None
        self.pre_centered = False

    @classmethod
    def from_native(cls, handle: ctypes.c_void_p, n: int, ids: np.ndarray,
                    nbytes: int) -> "ConformerLibrary":
        """A library whose rows live only in a native FssLibrary (see store.load_library)."""
        lib = cls.__new__(cls)
        lib.shape = lib.pharm = lib.shaep = lib.surface = None
        lib.n = int(n)
        lib.ids = np.asarray(ids)
        lib._native = handle
        lib._native_bytes = int(nbytes)
        lib._native_surface = False        # whether the native library holds surface rows
        lib.num_surf_points = None
        lib.pre_centered = False
        return lib

    @property
    def has_surface_rows(self) -> bool:
        """Whether the library holds surface rows (the ShaEP channel or surface rows alone)."""
        return (getattr(self, "shaep", None) is not None or getattr(self, "surface", None) is not None
                or bool(getattr(self, "_native_surface", False)))

    def id_of(self, index: int):
        """The id of library entry ``index`` as a Python scalar."""
        v = self.ids[index]
        return v.item() if isinstance(v, np.generic) else v

    def native(self) -> ctypes.c_void_p:
        """The library's copy in a native FssLibrary (created and uploaded on first use)."""
        if getattr(self, "_native", None) is None:
            lib = _lib.load()
            spec = _lib.FssLibSpec()
            some = next(c for c in (self.shape, self.pharm, self.shaep, self.surface) if c is not None)
            dev = (some.pts if isinstance(some, CloudSet) else
                   some.anc if isinstance(some, PharmSet) else some.srf.pts).device
            spec.device = dev.index if dev.index is not None else torch.cuda.current_device()
            spec.channels = (_lib.LIB_SHAPE if self.shape is not None else 0) | \
                (_lib.LIB_PHARM if self.pharm is not None else 0) | \
                (_lib.LIB_SHAEP if self.shaep is not None else 0) | \
                (_lib.LIB_SURFACE if self.surface is not None else 0)
            spec.max_conformers = max(self.n, 1)
            rows = _lib.FssLibRows()
            keep = []

            def counts(a):
                c = np.ascontiguousarray(a, np.int32)
                keep.append(c)
                return c.ctypes.data

            if self.shape is not None:
                spec.max_points = max(int(self.shape.counts.sum()), 1)
                rows.pts, rows.counts = _lib.ptr(self.shape.pts), counts(self.shape.counts)
            if self.pharm is not None:
                spec.max_pharm = max(int(self.pharm.counts.sum()), 1)
                rows.anc, rows.vec, rows.typ = (_lib.ptr(self.pharm.anc), _lib.ptr(self.pharm.vec),
                                                _lib.ptr(self.pharm.typ))
                rows.pcounts = counts(self.pharm.counts)
            if self.shaep is not None:
                se = self.shaep
                spec.max_surface = max(int(se.srf.counts.sum()), 1)
                spec.max_atoms = max(int(se.atm.counts.sum()), 1)
                rows.srf, rows.scounts = _lib.ptr(se.srf.pts), counts(se.srf.counts)
                rows.atm, rows.rad, rows.acounts = _lib.ptr(se.atm.pts), _lib.ptr(se.rad), counts(se.atm.counts)
            if self.surface is not None:
                spec.max_surface = max(int(self.surface.counts.sum()), 1)
                rows.srf, rows.scounts = _lib.ptr(self.surface.pts), counts(self.surface.counts)
            h = ctypes.c_void_p()
            _lib.call("fss_lib_create", _lib.byref(spec), ctypes.byref(h))
            torch.cuda.synchronize(dev)
            first = ctypes.c_int64()
            try:
                _lib.call("fss_lib_append", h, self.n, _lib.byref(rows), ctypes.byref(first))
            except Exception:
                lib.fss_lib_destroy(h)
                raise
            self._native = h
        return self._native

    def close(self) -> None:
        """Release the native copy (no screening service may still use it)."""
        if getattr(self, "_native", None) is not None:
            _lib.call("fss_lib_destroy", self._native)
            self._native = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_arrays(cls, points=None, esp=None, pharm_types=None, pharm_anchors=None,
                    pharm_vectors=None, ids=None, device="cuda", surf_points=None, surf_esp=None,
                    atoms_w_H=None, charges=None, radii=None) -> "ConformerLibrary":
        """Per-conformer arrays.  ``points`` [+ ``esp``]: the shape channel (heavy atoms for the
        ShaEP score at alpha 0.81); ``pharm_*``: pharmacophores; ``surf_points``, ``surf_esp``,
        ``atoms_w_H``, ``charges``, ``radii``: the ShaEP channel (vol_and_surf_esp), or with
        ``surf_points`` [+ ``surf_esp``] alone, surface rows for the surf / surf_esp modes."""
        shape = CloudSet.from_arrays(points, esp, device=device) if points is not None else None
        pharm = (PharmSet.from_arrays(pharm_types, pharm_anchors, pharm_vectors, device=device)
                 if pharm_types is not None else None)
        shaep = surface = None
        if surf_points is not None and atoms_w_H is not None:
            shaep = ShaepSet.from_arrays(surf_points, surf_esp, atoms_w_H, charges, radii, device=device)
        elif surf_points is not None:
            surface = CloudSet.from_arrays(surf_points, surf_esp, device=device)
        lib = cls(shape, pharm, ids, shaep, surface)
        surf = shaep.srf.counts if shaep is not None else surface.counts if surface is not None else None
        if surf is not None and len(surf) and (surf == surf[0]).all():
            lib.num_surf_points = int(surf[0])
        return lib

    def nbytes(self) -> int:
        if self._native_bytes is not None:
            return self._native_bytes
        tot = 0
        parts = [self.shape, self.pharm, self.surface]
        if self.shaep is not None:
            parts += [self.shaep.srf, self.shaep.atm, self.shaep]
        for cs in parts:
            if cs is None:
                continue
            for t in vars(cs).values():
                if isinstance(t, torch.Tensor):
                    tot += t.numel() * t.element_size()
        return tot


def _surf_count(library) -> Optional[int]:
    """Surface points per conformer: the library's record, else the common size of its shape
    clouds (surface clouds in the surf modes)."""
    if getattr(library, "num_surf_points", None):
        return int(library.num_surf_points)
    sh = getattr(library, "shape", None)
    if sh is not None and len(sh.counts) and (sh.counts == sh.counts[0]).all():
        return int(sh.counts[0])
    return None


def resolve_alpha(library, mode: str, alpha: Optional[float]) -> float:
    """The shape width a screen uses (screen.py:936-959): ``ALPHA(num_surf_points)`` for surf /
    surf_esp, 0.81 for the atom modes, and vol_and_surf_esp needs it given (0.81 selects the
    heavy-atom shape term, anything else the surface one)."""
    from .constants import ALPHA
    if alpha is not None:
        return float(alpha)
    if mode in ("surf", "surf_esp"):
        n = _surf_count(library)
        if n is None:
            raise ValueError("surf / surf_esp screening takes alpha = ALPHA(num_surf_points), but the "
                             "library's surface clouds differ in size: pass alpha=...")
        return float(ALPHA(n))
    if mode == "vol_and_surf_esp":
        raise ValueError("vol_and_surf_esp requires an explicit alpha (it selects volumetric shape at "
                         "alpha=0.81, otherwise surface shape); pass alpha=...")
    return 0.81


def plan_launch(remaining: Sequence[int], capacity: int) -> List[int]:
    """Fill ``capacity`` library entries from requests given oldest first, with ``remaining``
    entries each: every request takes what it needs of what the earlier ones left (FIFO
    filling; the tail of one query and the head of the next share a launch).  Host restatement
    of the native planner ``fss_plan_launch``."""
    take = []
    left = max(int(capacity), 0)
    for r in remaining:
        k = min(max(int(r), 0), left)
        take.append(k)
        left -= k
    return take


def query_row_bytes(max_query_points: int, max_query_pharm: int, max_query_surface: int = 0,
                    max_query_atoms: int = 0) -> int:
    """Device bytes of one query of the maximum size in the service's row pools."""
    counts = (max_query_points, max_query_pharm, max_query_surface, max_query_atoms)
    return sum(int(n) * b for n, b in zip(counts, _lib.QPOOL_ROW_BYTES))


def arena_bytes(launch_pairs: int, restarts: int, max_active: int, top_k: int,
                max_query_points: int, max_query_pharm: int, max_query_surface: int = 0,
                max_query_atoms: int = 0, query_budget_bytes: int = 0) -> int:
    """Device bytes a service reserves at creation (fss_screen_create's arena), at most: the
    launch arena, ``max_active`` top-k slots and the query row pools (``query_budget_bytes``, or
    ``max_active`` queries of the maximum size when that is 0 or larger)."""
    P, B, A, K = launch_pairs, restarts, max_active, top_k
    per_pair = B * (4 + 3) * 4 + B * 8 + 8 + 4 + 64 + 4 + 32 + 64
    per_slot = 2 * K * (8 + 8 + 64) + 64
    rows = A * query_row_bytes(max_query_points, max_query_pharm, max_query_surface, max_query_atoms)
    if query_budget_bytes:
        rows = min(rows, int(query_budget_bytes))
    return P * per_pair + A * per_slot + rows + B * 16


def plan_query_pools(mode: int, max_active: int, max_query_points: int = 0, max_query_pharm: int = 0,
                     max_query_surface: int = 0, max_query_atoms: int = 0,
                     query_budget_bytes: int = 0, shape_from_surface: bool = False) -> List[int]:
    """Rows of the four query pools a service with these parameters reserves (native
    ``fss_plan_query_pools``; raises when the budget does not hold one query of the maximum
    size)."""
    sp = _lib.FssScoreParams()
    sp.mode = int(mode)
    prm = _lib.FssScreenParams()
    prm.max_active = int(max_active)
    prm.max_query_points, prm.max_query_pharm = int(max_query_points), int(max_query_pharm)
    prm.max_query_surface, prm.max_query_atoms = int(max_query_surface), int(max_query_atoms)
    prm.query_budget_bytes, prm.shape_from_surface = int(query_budget_bytes), int(shape_from_surface)
    rows = (ctypes.c_int64 * _lib.QPOOLS)()
    _lib.call("fss_plan_query_pools", _lib.byref(sp), _lib.byref(prm), rows)
    return [int(x) for x in rows]


class RowPool:
    """The service's first-fit pool of device rows, host side only (native ``fss_rowpool_*``):
    contiguous ranges handed out first-fit and merged with free neighbours on release."""

    def __init__(self, rows: int):
        self._h = ctypes.c_void_p()
        _lib.call("fss_rowpool_create", ctypes.c_int64(int(rows)), ctypes.byref(self._h))

    def alloc(self, n: int) -> Optional[int]:
        """First row of a free range of ``n`` rows, or None when none is long enough."""
        first = ctypes.c_int64()
        rc = _lib.load().fss_rowpool_alloc(self._h, int(n), ctypes.byref(first))
        if rc == _lib.PENDING:
            return None
        if rc:
            raise _lib.FssError("fss_rowpool_alloc", rc, _lib.load().fss_last_error().decode())
        return int(first.value)

    def free(self, first: int, n: int) -> None:
        _lib.call("fss_rowpool_free", self._h, int(first), int(n))

    def stats(self) -> tuple:
        """(free rows, longest free range)."""
        f, l = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("fss_rowpool_stats", self._h, ctypes.byref(f), ctypes.byref(l))
        return int(f.value), int(l.value)

    def close(self) -> None:
        if self._h:
            _lib.call("fss_rowpool_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def admit_fifo(pools: Sequence[RowPool], rows: Sequence[Sequence[int]], free_slots: int) -> List[List[int]]:
    """The service's admission step (native ``fss_admit_fifo``): queries given oldest first,
    needing ``rows[k][c]`` rows of ``pools[c]``, are admitted in order while slots are free and
    every pool has room; the first that does not fit stops the step.  Returns the admitted
    queries' first rows, one list per query."""
    n, npools = len(rows), len(pools)
    flat = (ctypes.c_int64 * max(1, n * npools))(*[int(x) for r in rows for x in r])
    first = (ctypes.c_int64 * max(1, n * npools))()
    handles = (ctypes.c_void_p * max(1, npools))(*[p._h.value for p in pools])
    got = ctypes.c_int32()
    _lib.call("fss_admit_fifo", handles, npools, n, flat, int(free_slots), first, ctypes.byref(got))
    return [[int(first[k * npools + c]) for c in range(npools)] for k in range(got.value)]


@dataclass
class ScreenResult:
    hits: List[Hit]
    latency_ms: float
    queue_ms: float
    pairs: int


class ScreenService:
    """Serve screening queries against a resident library with continuous batching (a thin
    host wrapper over the native ``fss_screen_*`` service).

    The alignment keywords follow the reference screen (screen.py:851-882): ``num_repeats`` and
    ``max_num_steps`` default to the mode's table, ``lam`` is the raw lambda (scaled by
    LAM_SCALING for surf_esp and vol_and_surf_esp, defaults 0.3 / 0.1 / 0.001), and
    vol_and_surf_esp scores heavy atoms at alpha 0.81 and surface points otherwise.
    ``extended_points`` / ``only_extended`` apply to the directional pharmacophore channel
    (pharm, or vol_color with ``directionless=False``), as in screen.py:866-867.
    Tiling (``lanes_per_pose``, ``pairs_per_wg``, ``variant``, ``groups_per_wg``: fss_align's
    meanings) defaults by mode: atom clouds 2 lanes per restart and 32 / restarts pairs per
    workgroup; surf / surf_esp the split-bf16 MFMA loop with two wavefronts per pair; ShaEP 32
    lanes per restart.
    """

    def __init__(self, library: ConformerLibrary, mode: str = "vol", *, num_repeats: Optional[int] = None,
                 max_num_steps: Optional[int] = None, lr: float = 0.1, top_k: int = 100,
                 launch_pairs: int = 65536, hbm_budget_bytes: int = 8 << 30, max_active: int = 16,
                 query_budget_bytes: Optional[int] = None,
                 early_stop: bool = True, alpha: Optional[float] = None, lam: Optional[float] = None,
                 similarity: str = "tanimoto", directionless: bool = True,
                 extended_points: bool = False, only_extended: bool = False,
                 lanes_per_pose: Optional[int] = None, pairs_per_wg: Optional[int] = None,
                 variant: Optional[int] = None, groups_per_wg: Optional[int] = None,
                 max_query_points: int = 4096,
                 max_query_pharm: int = 256, esp_weight: float = 0.5, probe_radius: float = 1.0,
                 max_query_surface: int = 4096, max_query_atoms: int = 1024, device=None):
        from .modes import MODE_SEEDS, MODE_STEPS, canonical
        mode = canonical(mode)
        if mode not in _KIND:
            raise ValueError(f"unsupported screening mode {mode!r}")
        alpha = resolve_alpha(library, mode, alpha)
        if not 1 <= int(top_k) <= _lib.SCREEN_MAX_TOPK:
            raise ValueError(f"top_k must be in [1, {_lib.SCREEN_MAX_TOPK}]")
        self.lib, self.mode = library, mode
        combo = mode == "vol_and_surf_esp"
        if mode == "surf_esp":
            lam = LAM_SCALING * (0.3 if lam is None else lam)
        elif combo:
            lam = LAM_SCALING * (0.001 if lam is None else lam)
        elif lam is None:
            lam = 0.1
        self.spec = ScoreSpec(mode=_KIND[mode], alpha=alpha, lam=lam, similarity=similarity,
                              directionless=directionless,
                              extended_points=bool(extended_points) and not directionless,
                              only_extended=bool(only_extended) and not directionless,
                              esp_weight=esp_weight,
                              probe_radius=probe_radius)
        self.B = int(MODE_SEEDS[mode] if num_repeats is None else num_repeats)
        steps = MODE_STEPS[mode] if max_num_steps is None else max_num_steps
        self.opt = OptimSpec(max_steps=int(steps), lr=float(lr), early_stop=bool(early_stop))
        self.top_k = int(top_k)
        self.launch_pairs = int(launch_pairs)
        # vol_and_surf_esp: the shape term takes the heavy atoms at alpha 0.81, else the surface
        # rows of the ShaEP channel (_core.py:1079-1080)
        # surf / surf_esp: the library's surface rows when it has them, else its shape channel
        self.surface_mode = mode in ("surf", "surf_esp") and library.has_surface_rows
        self.shape_from_surface = (combo and alpha != 0.81) or self.surface_mode
        shape = self.spec.mode != _lib.MODE_PHARM and not self.shape_from_surface
        pharm = self.spec.mode in (_lib.MODE_PHARM, _lib.MODE_SHAPE_PHARM)
        mqp, mqh = (int(max_query_points) if shape else 0), (int(max_query_pharm) if pharm else 0)
        mqs = int(max_query_surface) if (combo or self.surface_mode) else 0
        mqa = int(max_query_atoms) if combo else 0
        # admission by memory: the launch arena, then up to max_active top-k slots, then the query
        # row pools from what is left of the budget (query_budget_bytes overrides that share)
        one = query_row_bytes(mqp, mqh, mqs, mqa)
        fixed = arena_bytes(self.launch_pairs, self.B, 0, self.top_k, 0, 0)
        slot = arena_bytes(0, 0, 1, self.top_k, 0, 0)
        if fixed + slot + one > hbm_budget_bytes:
            raise MemoryError("the HBM budget does not hold one launch, one query slot and one query")
        self.max_active = max(1, min(int(max_active), (hbm_budget_bytes - fixed - one) // slot,
                                     self.launch_pairs))
        if query_budget_bytes is None:
            query_budget_bytes = hbm_budget_bytes - fixed - self.max_active * slot
        if query_budget_bytes < one:
            raise MemoryError(f"query_budget_bytes {query_budget_bytes} does not hold one query of the "
                              f"maximum size ({one} bytes)")
        self.query_budget_bytes = int(min(query_budget_bytes, self.max_active * one))
        self.table = restart_table(self.B)
        if lanes_per_pose is None and pairs_per_wg is None:
            # ShaEP poses carry the surface-atom ESP sums: a wider tile.  Atom-sized clouds: 2 lanes
            # per restart and 32 / B pairs per workgroup, so the per-step pose work is shared by
            # several pairs' restarts (10 restarts: 3 pairs x 10 x 2 lanes).  Surfaces: every restart
            # of a pair in one wavefront with as many lanes each as fit (10 restarts: 6 lanes).
            if not combo and mode not in ("surf", "surf_esp") and 32 // self.B >= 2:
                lanes_per_pose, pairs_per_wg = 2, 32 // self.B
            elif mode in ("surf", "surf_esp") and variant is None and groups_per_wg is None:
                # surfaces: the split-bf16 MFMA loop, a wavefront per pose, two wavefronts per
                # pair taking the restarts in turn (1M-entry-sized sweep, 100k conformers of 200
                # points + ESP, 10 x 100: 100k pairs/s on 6-lane VALU groups, 85k / 81k / 117k /
                # 123k on the MFMA loop with 10 / 5 / 4 / 2 wavefronts)
                lanes_per_pose, variant, groups_per_wg = 64, 3, min(2, self.B)
            else:
                lanes_per_pose = 32 if combo else max(1, 64 // self.B)
        elif lanes_per_pose is None:
            lanes_per_pose = 32 if combo else max(1, 64 // (self.B * int(pairs_per_wg)))
        self.lanes_per_pose, self.pairs_per_wg = int(lanes_per_pose), int(pairs_per_wg or 1)
        self.variant, self.groups_per_wg = int(variant or 0), int(groups_per_wg or 0)
        prm = _lib.FssScreenParams()
        prm.n_restarts, prm.max_steps = self.B, self.opt.max_steps
        prm.lr, prm.beta1, prm.beta2, prm.eps = self.opt.lr, self.opt.beta1, self.opt.beta2, self.opt.eps
        prm.early_stop, prm.es_tol = int(self.opt.early_stop), self.opt.es_tol
        prm.es_patience = int(self.opt.es_patience)
        prm.top_k, prm.launch_pairs, prm.max_active = self.top_k, self.launch_pairs, self.max_active
        prm.max_query_points, prm.max_query_pharm = mqp, mqh
        prm.max_query_surface, prm.max_query_atoms = mqs, mqa
        prm.shape_from_surface = int(self.shape_from_surface)
        prm.query_budget_bytes = self.query_budget_bytes
        prm.lanes_per_pose, prm.restarts_per_wg, prm.variant = self.lanes_per_pose, self.B, self.variant
        prm.pairs_per_wg = self.pairs_per_wg
        prm.groups_per_wg = self.groups_per_wg
        prm.seed_flags = (_lib.SEED_PHARM if self.spec.mode == _lib.MODE_PHARM else
                          _lib.SEED_SURF if combo else 0)
        prm.quat_table = self.table.ctypes.data
        self._sp = self.spec.struct()
        self._h = ctypes.c_void_p()
        _lib.call("fss_screen_create", library.native(), _lib.byref(self._sp), _lib.byref(prm),
                  ctypes.byref(self._h))
        self._futures: Dict[int, Future] = {}
        self._submitted: Dict[int, int] = {}
        self._outs: Dict[int, object] = {}
        self._cv = threading.Condition()
        self._stop = False
        self._poller = threading.Thread(target=self._poll_loop, name="fss-screen-poll", daemon=True)
        self._poller.start()

    # -- public ------------------------------------------------------------------------------
    def submit(self, query, top_k: Optional[int] = None, scores_out=None) -> Future:
        """Queue a query and return a Future resolving to a ScreenResult.

        ``query`` is a dict with the arrays its mode needs: ``points`` [+ ``esp``] and/or
        ``pharm_types``, ``pharm_anchors``, ``pharm_vectors``; vol_and_surf_esp: ``surf_points``,
        ``surf_esp``, ``atoms_w_H``, ``charges``, ``radii`` and, at alpha 0.81, ``points`` = the
        heavy atoms.  An object with the container's ``Molecule`` / ``MoleculeProfile``
        attributes is taken too (see :func:`query_arrays`).  ``scores_out``: an array of at
        least the library's size, written with every entry's best score in library order
        before the future resolves (screen.py:1174-1176)."""
        query = query_arrays(query, self.mode)
        k = int(top_k or self.top_k)
        if not 1 <= k <= self.top_k:
            raise ValueError(f"top_k must be in [1, {self.top_k}] for this service")
        if scores_out is not None and len(scores_out) < self.lib.n:
            raise ValueError(f"scores_out holds {len(scores_out)} entries, the library {self.lib.n}")
        q = _lib.FssQuery()
        q.all_scores = int(scores_out is not None)
        keep = []

        def rows4(xyz, w=None):
            xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
            r = np.zeros((len(xyz), 4), np.float32)
            r[:, :3] = xyz
            if w is not None:
                r[:, 3] = np.asarray(w, np.float32).reshape(-1)
            keep.append(r)
            return r

        if self.spec.mode == _lib.MODE_ESP_COMBO:
            srf = rows4(query["surf_points"], query["surf_esp"])
            atm = rows4(query["atoms_w_H"], query["charges"])
            rad = np.ascontiguousarray(query["radii"], np.float32).reshape(-1)
            if len(rad) != len(atm):
                raise ValueError("radii must have one entry per atom")
            keep.append(rad)
            q.n_surface, q.srf = len(srf), srf.ctypes.data
            q.n_atoms, q.atm, q.rad = len(atm), atm.ctypes.data, rad.ctypes.data
        if self.surface_mode:      # the query's surface: points [+ esp] as surface rows
            srf = rows4(query["points"], query.get("esp") if self.spec.mode == _lib.MODE_SHAPE_ESP else None)
            q.n_surface, q.srf = len(srf), srf.ctypes.data
        if self.spec.mode != _lib.MODE_PHARM and not self.shape_from_surface:
            xyz = np.asarray(query["points"], np.float32).reshape(-1, 3)
            rows = np.zeros((len(xyz), 4), np.float32)
            rows[:, :3] = xyz
            if self.spec.mode == _lib.MODE_SHAPE_ESP:
                rows[:, 3] = np.asarray(query["esp"], np.float32).reshape(-1)
            q.n_points, q.pts = len(rows), rows.ctypes.data
            keep.append(rows)
        if self.spec.mode in (_lib.MODE_PHARM, _lib.MODE_SHAPE_PHARM):
            typ = np.ascontiguousarray(query["pharm_types"], np.int32).reshape(-1)
            anc = np.zeros((len(typ), 4), np.float32)
            vec = np.zeros((len(typ), 4), np.float32)
            anc[:, :3] = np.asarray(query["pharm_anchors"], np.float32).reshape(-1, 3)
            vec[:, :3] = np.asarray(query["pharm_vectors"], np.float32).reshape(-1, 3)
            q.n_pharm, q.anc, q.vec, q.typ = len(typ), anc.ctypes.data, vec.ctypes.data, typ.ctypes.data
            keep += [typ, anc, vec]
        fut: Future = Future()
        ticket = ctypes.c_int64()
        with self._cv:
            if self._stop:
                raise RuntimeError("the screening service is closed")
            _lib.call("fss_screen_submit", self._h, ctypes.byref(q), ctypes.byref(ticket))
            self._futures[ticket.value] = fut
            self._submitted[ticket.value] = k
            if scores_out is not None:
                self._outs[ticket.value] = scores_out
            self._cv.notify_all()
        return fut

    @property
    def launches(self) -> int:
        """Launches run so far (the final count once closed)."""
        with self._cv:
            if self._stop:
                return self._final_launches
            n = ctypes.c_int64()
            _lib.call("fss_screen_stats", self._h, ctypes.byref(n), None, None)
            return int(n.value)

    def stats(self) -> dict:
        """Launches run, (query, library entry) pairs aligned, optimiser steps those pairs ran
        (summed over pairs) and the row-pair terms those steps evaluated (steps x restarts x
        query rows x entry rows, per score term), so far (the final counts once closed)."""
        with self._cv:
            if self._stop:
                return dict(self._final_stats)
            v = [ctypes.c_int64() for _ in range(3)]
            _lib.call("fss_screen_stats", self._h, *(ctypes.byref(x) for x in v))
            w = [ctypes.c_double() for _ in range(3)]
            _lib.call("fss_screen_terms", self._h, *(ctypes.byref(x) for x in w))
            return {"launches": int(v[0].value), "pairs": int(v[1].value), "pair_steps": int(v[2].value),
                    "shape_terms": w[0].value, "pharm_terms": w[1].value, "esp_terms": w[2].value}

    def admission(self) -> dict:
        """Rows of the four query pools, the most queries admitted at once, and how many queries
        waited for pool room while a slot was free."""
        rows = (ctypes.c_int64 * _lib.QPOOLS)()
        most, waits = ctypes.c_int64(), ctypes.c_int64()
        with self._cv:
            if self._stop:
                return dict(self._final_admission)
            _lib.call("fss_screen_admission", self._h, rows, ctypes.byref(most), ctypes.byref(waits))
        return {"pool_rows": [int(x) for x in rows], "max_admitted": int(most.value),
                "memory_waits": int(waits.value)}

    def close(self) -> None:
        final = self.stats()
        final_adm = self.admission() if not self._stop else None
        with self._cv:
            if self._stop:
                return
            n = ctypes.c_int64()
            _lib.call("fss_screen_stats", self._h, ctypes.byref(n), None, None)
            self._final_launches = int(n.value)
            self._final_stats = final
            self._final_admission = final_adm
            self._stop = True
            self._cv.notify_all()
        self._poller.join()
        _lib.call("fss_screen_destroy", self._h)
        for fut in self._futures.values():
            if not fut.done():
                fut.set_exception(RuntimeError("the screening service was closed"))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- result delivery ----------------------------------------------------------------------
    def _poll_loop(self) -> None:
        lib = _lib.load()
        K = self.top_k
        score = np.zeros(K, np.float64)
        index = np.zeros(K, np.int64)
        tf = np.zeros((K, 16), np.float32)
        hits = _lib.FssHits()
        hits.capacity = K
        hits.score, hits.index, hits.transform = score.ctypes.data, index.ctypes.data, tf.ctypes.data
        all_buf = None
        while True:
            with self._cv:
                while not self._futures and not self._stop:
                    self._cv.wait()
                if self._stop:
                    return
                tickets = sorted(self._futures)
                outs = {t: self._outs.get(t) for t in tickets}
            # wait (in native code, GIL released) up to 5 ms on the oldest ticket, then sweep
            for i, t in enumerate(tickets):
                out = outs[t]
                if out is not None:
                    if all_buf is None or len(all_buf) < self.lib.n:
                        all_buf = np.empty(self.lib.n, np.float64)
                    hits.all_scores, hits.n_all = all_buf.ctypes.data, len(all_buf)
                else:
                    hits.all_scores, hits.n_all = None, 0
                rc = lib.fss_screen_poll(self._h, t, 5 if i == 0 else 0, ctypes.byref(hits))
                if rc == _lib.PENDING:
                    continue
                with self._cv:
                    fut = self._futures.pop(t)
                    k = self._submitted.pop(t)
                    self._outs.pop(t, None)
                if rc != 0:
                    fut.set_exception(_lib.FssError("fss_screen_poll", rc, lib.fss_last_error().decode()))
                    continue
                if out is not None:
                    out[:int(hits.n_all)] = all_buf[:int(hits.n_all)]
                n = min(int(hits.n), k)
                res = [Hit(float(score[i]), self.lib.id_of(int(index[i])), tf[i].reshape(4, 4).copy())
                       for i in range(n)]
                fut.set_result(ScreenResult(res, float(hits.latency_ms), float(hits.queue_ms), self.lib.n))


def _as_library(library, mode: str, device, center: bool = False):
    """(ConformerLibrary, owned): a ProfileStore or store path is loaded into HBM first, its
    conformers centred on their heavy-atom means when ``center`` and the store is not
    pre-centred."""
    from .store import ProfileStore, load_library
    if isinstance(library, (ProfileStore, str, os.PathLike)):
        return load_library(library, mode, device=device if device is not None else 0, center=center), True
    return library, False


def query_arrays(query, mode: str) -> dict:
    """The query dict of :meth:`ScreenService.submit` from a dict (returned as is) or from an
    object with the container's ``Molecule`` / ``MoleculePair`` attributes (``atom_pos``,
    ``surf_pos``, ``surf_esp``, ``pharm_types``, ``pharm_ancs``, ``pharm_vecs``, heavy-atom
    ``atom_charges`` or ``partial_charges`` [+ ``nonH_atoms_idx``], ``centers_w_H``, ``radii``)."""
    if isinstance(query, dict):
        return query
    g = lambda k: getattr(query, k, None)
    d = {"atom_pos": g("atom_pos")}
    if mode in ("vol", "vol_color", "vol_and_surf_esp"):
        d["points"] = g("atom_pos")
    if mode == "vol_esp":
        ch = g("atom_charges")
        if ch is None and g("partial_charges") is not None:
            idx = g("nonH_atoms_idx")
            ch = np.asarray(g("partial_charges"))[idx] if idx is not None else g("partial_charges")
        d["points"], d["esp"] = g("atom_pos"), ch
    if mode in ("surf", "surf_esp"):
        d["points"] = g("surf_pos")
        if mode == "surf_esp":
            d["esp"] = g("surf_esp")
    if mode in ("pharm", "vol_color"):
        d.update(pharm_types=g("pharm_types"), pharm_anchors=g("pharm_ancs"), pharm_vectors=g("pharm_vecs"))
    if mode == "vol_and_surf_esp":
        d.update(surf_points=g("surf_pos"), surf_esp=g("surf_esp"), atoms_w_H=g("centers_w_H"),
                 charges=g("partial_charges"), radii=g("radii"))
    missing = [k for k, v in d.items() if v is None and k != "atom_pos"]
    if missing:
        raise ValueError(f"{mode} query is missing {missing}")
    return d


_COORDS = ("points", "surf_points", "atoms_w_H", "pharm_anchors", "atom_pos")


def centered_query(query: dict, mode: str) -> dict:
    """A copy of the query dict moved so that its heavy atoms' mean is the origin (``atom_pos``
    when present, else the mean of the rows the restarts are seeded from); vectors and ESP
    values are unchanged.  The caller's arrays are not modified."""
    if query.get("atom_pos") is not None:
        c = np.asarray(query["atom_pos"], np.float64).reshape(-1, 3).mean(0)
    elif mode == "pharm":
        c = np.asarray(query["pharm_anchors"], np.float64).reshape(-1, 3).mean(0)
    elif mode == "vol_and_surf_esp":
        c = np.asarray(query["surf_points"], np.float64).reshape(-1, 3).mean(0)
    else:
        c = np.asarray(query["points"], np.float64).reshape(-1, 3).mean(0)
    out = dict(query)
    for k in _COORDS:
        if out.get(k) is not None:
            out[k] = (np.asarray(out[k], np.float64).reshape(-1, 3) - c).astype(np.float32)
    return out


def screen(query, library, mode: str = "surf_esp", *, backend: Optional[str] = None,
           do_center: Optional[bool] = None, top_k: int = 1000, ndev: Optional[int] = None,
           scores_out=None, alpha: Optional[float] = None, progress: bool = False, device=None,
           **align_kwargs) -> List[Hit]:
    """One query against the whole library (a ConformerLibrary, a ProfileStore or a store path);
    ``Hit(score, id, transform)`` sorted descending.  The reference's signature
    (screen.py:1141-1192); see :func:`screen_many`."""
    outs = None if scores_out is None else [scores_out]
    return screen_many([query], library, mode, backend=backend, do_center=do_center, top_k=top_k,
                       ndev=ndev, scores_out=outs, alpha=alpha, progress=progress, device=device,
                       **align_kwargs)[0]


def screen_many(queries: Sequence, library, mode: str = "surf_esp", *, backend: Optional[str] = None,
                do_center: Optional[bool] = None, top_k: int = 1000, ndev: Optional[int] = None,
                scores_out=None, alpha: Optional[float] = None, progress: bool = False, device=None,
                **align_kwargs) -> List[List[Hit]]:
    """A panel of queries served together (continuous batching); one hit list per query
    (screen.py:1038-1138).

    * ``backend``: None or a HIP alias; ``ndev``: one GPU (None or 1).
    * ``alpha``: ``ALPHA(num_surf_points)`` for surf / surf_esp, 0.81 for the atom modes,
      required for vol_and_surf_esp; vol_esp requires ``lam`` (screen.py:936-959).
    * ``do_center`` (default on): queries are centred on their heavy-atom mean; a store that
      is not pre-centred is centred on load.  A ConformerLibrary passed in is used as it is.
      The transforms move library entries onto the (centred) query.
    * ``scores_out``: one array per query of at least the library's size, written with every
      entry's best score in library order.
    * ``progress`` is accepted; the service reports no progress.
    """
    from .container import resolve_backend
    from .modes import canonical
    resolve_backend(backend)
    if ndev not in (None, 1):
        raise ValueError("screening runs on one GPU here (ndev=1)")
    mode = canonical(mode)
    if align_kwargs.pop("no_H", True) is False:
        raise ValueError("screen() aligns heavy atoms only; no_H=False is not supported")
    if mode == "vol_esp" and "lam" not in align_kwargs:
        raise ValueError("vol_esp requires an explicit lam=... (the ESP/partial-charge weight); pass lam=...")
    queries = list(queries)
    if scores_out is not None and len(scores_out) != len(queries):
        raise ValueError("scores_out needs one array per query")
    if mode == "vol_and_surf_esp":
        resolve_alpha(None, mode, alpha)          # needs no library: fail before loading one
    center = True if do_center is None else bool(do_center)
    lib, owned = _as_library(library, mode, device, center=center)
    try:
        alpha = resolve_alpha(lib, mode, alpha)
        qs = [query_arrays(q, mode) for q in queries]
        if center:
            qs = [centered_query(q, mode) for q in qs]
        outs = scores_out if scores_out is not None else [None] * len(qs)
        with ScreenService(lib, mode, top_k=top_k, alpha=alpha, **align_kwargs) as svc:
            futs = [svc.submit(q, scores_out=o) for q, o in zip(qs, outs)]
            return [f.result().hits for f in futs]
    finally:
        if owned:
            lib.close()
