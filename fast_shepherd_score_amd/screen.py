"""Screening: queries against a conformer library resident in HBM, with continuous batching.

Front-end restating the reference's screening contract (shepherd_score/screen.py:74, 1038-1192):
``screen(query, library, mode, top_k, **align_kwargs) -> [Hit(score, id, transform)]`` sorted by
score, descending.  The reference streams ``ProfileStore`` shards from disk through per-query
aligner calls; here the library is packed once into HBM and served by the native screening
service of libfss_hip.so (``fss_lib_*`` / ``fss_screen_*`` in include/fss.h), which this module
wraps:

* admission: a query takes one of ``max_active`` slots of a device arena reserved when the
  service is created (sized from the HBM budget); the rest wait in FIFO order;
* continuous batching: every launch is filled from *all* admitted queries (water-filling over
  their remaining library ranges, :func:`plan_launch` / ``fss_plan_launch``), so a query that
  arrives while others are on the GPU joins the very next launch;
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
from .packing import CloudSet, PharmSet
from .seeds import quats_from_fibo

Hit = collections.namedtuple("Hit", ["score", "id", "transform"])

_KIND = {"vol": _lib.MODE_SHAPE, "surf": _lib.MODE_SHAPE, "vol_esp": _lib.MODE_SHAPE_ESP,
         "surf_esp": _lib.MODE_SHAPE_ESP, "pharm": _lib.MODE_PHARM, "vol_color": _lib.MODE_SHAPE_PHARM}


def restart_table(B: int) -> np.ndarray:
    """Screening restarts: identity, then rotations onto B-1 Fibonacci-sphere points."""
    q = np.zeros((B, 4), np.float32)
    q[0, 0] = 1.0
    if B > 1:
        q[1:] = quats_from_fibo(B - 1).astype(np.float32)
    return q


class ConformerLibrary:
    """Packed conformers on one device: shape channel (atoms or surface points [+ ESP]) and/or
    pharmacophores, with integer ids."""

    def __init__(self, shape: Optional[CloudSet] = None, pharm: Optional[PharmSet] = None,
                 ids: Optional[np.ndarray] = None):
        self.shape, self.pharm = shape, pharm
        n = shape.n if shape is not None else pharm.n
        self.n = n
        self.ids = np.arange(n, dtype=np.int64) if ids is None else np.asarray(ids)
        self._native = None
        self._native_bytes = None

    @classmethod
    def from_native(cls, handle: ctypes.c_void_p, n: int, ids: np.ndarray,
                    nbytes: int) -> "ConformerLibrary":
        """A library whose rows live only in a native FssLibrary (see store.load_library)."""
        lib = cls.__new__(cls)
        lib.shape = lib.pharm = None
        lib.n = int(n)
        lib.ids = np.asarray(ids)
        lib._native = handle
        lib._native_bytes = int(nbytes)
        return lib

    def id_of(self, index: int):
        """The id of library entry ``index`` as a Python scalar."""
        v = self.ids[index]
        return v.item() if isinstance(v, np.generic) else v

    def native(self) -> ctypes.c_void_p:
        """The library's copy in a native FssLibrary (created and uploaded on first use)."""
        if getattr(self, "_native", None) is None:
            lib = _lib.load()
            spec = _lib.FssLibSpec()
            dev = (self.shape.pts if self.shape is not None else self.pharm.anc).device
            spec.device = dev.index if dev.index is not None else torch.cuda.current_device()
            spec.channels = (_lib.LIB_SHAPE if self.shape is not None else 0) | \
                (_lib.LIB_PHARM if self.pharm is not None else 0)
            spec.max_conformers = max(self.n, 1)
            spec.max_points = max(int(self.shape.counts.sum()), 1) if self.shape is not None else 0
            spec.max_pharm = max(int(self.pharm.counts.sum()), 1) if self.pharm is not None else 0
            h = ctypes.c_void_p()
            _lib.call("fss_lib_create", _lib.byref(spec), ctypes.byref(h))
            cnt = pcnt = None
            args = [None] * 6
            if self.shape is not None:
                cnt = np.ascontiguousarray(self.shape.counts, np.int32)
                args[0], args[1] = _lib.ptr(self.shape.pts), cnt.ctypes.data
            if self.pharm is not None:
                pcnt = np.ascontiguousarray(self.pharm.counts, np.int32)
                args[2:] = [_lib.ptr(self.pharm.anc), _lib.ptr(self.pharm.vec),
                            _lib.ptr(self.pharm.typ), pcnt.ctypes.data]
            torch.cuda.synchronize(dev)
            first = ctypes.c_int64()
            try:
                _lib.call("fss_lib_upload", h, self.n, *args, ctypes.byref(first))
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
                    pharm_vectors=None, ids=None, device="cuda") -> "ConformerLibrary":
        shape = CloudSet.from_arrays(points, esp, device=device) if points is not None else None
        pharm = (PharmSet.from_arrays(pharm_types, pharm_anchors, pharm_vectors, device=device)
                 if pharm_types is not None else None)
        return cls(shape, pharm, ids)

    def nbytes(self) -> int:
        if self._native_bytes is not None:
            return self._native_bytes
        tot = 0
        for cs in (self.shape, self.pharm):
            if cs is None:
                continue
            for t in vars(cs).values():
                if isinstance(t, torch.Tensor):
                    tot += t.numel() * t.element_size()
        return tot


def plan_launch(remaining: Sequence[int], capacity: int) -> List[int]:
    """Split ``capacity`` library entries over requests with ``remaining`` entries each:
    water-filling, so every active request advances in every launch (continuous batching).
    Host restatement of the native planner ``fss_plan_launch``."""
    take = [0] * len(remaining)
    left = capacity
    live = [i for i, r in enumerate(remaining) if r > 0]
    while left > 0 and live:
        share = max(1, left // len(live))
        nxt = []
        for i in live:
            k = min(share, remaining[i] - take[i], left)
            take[i] += k
            left -= k
            if take[i] < remaining[i]:
                nxt.append(i)
            if left == 0:
                break
        live = nxt
    return take


def arena_bytes(launch_pairs: int, restarts: int, max_active: int, top_k: int,
                max_query_points: int, max_query_pharm: int) -> int:
    """Device bytes a service reserves at creation (fss_screen_create's arena)."""
    P, B, A, K = launch_pairs, restarts, max_active, top_k
    per_pair = B * (4 + 3) * 4 + B * 8 + 8 + 4 + 64 + 4 + 32 + 32
    per_slot = max_query_points * 16 + max_query_pharm * 36 + 2 * K * (8 + 8 + 64) + 64
    return P * per_pair + A * per_slot + B * 16


@dataclass
class ScreenResult:
    hits: List[Hit]
    latency_ms: float
    queue_ms: float
    pairs: int


class ScreenService:
    """Serve screening queries against a resident library with continuous batching (a thin
    host wrapper over the native ``fss_screen_*`` service)."""

    def __init__(self, library: ConformerLibrary, mode: str = "vol", *, num_repeats: int = 10,
                 max_num_steps: int = 100, lr: float = 0.1, top_k: int = 100,
                 launch_pairs: int = 65536, hbm_budget_bytes: int = 8 << 30, max_active: int = 16,
                 early_stop: bool = True, alpha: float = 0.81, lam: Optional[float] = None,
                 similarity: str = "tanimoto", directionless: bool = True,
                 lanes_per_pose: int = 4, max_query_points: int = 4096,
                 max_query_pharm: int = 256, device=None):
        if mode not in _KIND:
            raise ValueError(f"unsupported screening mode {mode!r}")
        if not 1 <= int(top_k) <= _lib.SCREEN_MAX_TOPK:
            raise ValueError(f"top_k must be in [1, {_lib.SCREEN_MAX_TOPK}]")
        self.lib, self.mode = library, mode
        if lam is None:
            lam = 0.3 * LAM_SCALING if mode == "surf_esp" else 0.1
        self.spec = ScoreSpec(mode=_KIND[mode], alpha=alpha, lam=lam, similarity=similarity,
                              directionless=directionless)
        self.B = int(num_repeats)
        self.opt = OptimSpec(max_steps=int(max_num_steps), lr=float(lr), early_stop=bool(early_stop))
        self.top_k = int(top_k)
        self.launch_pairs = int(launch_pairs)
        shape = self.spec.mode != _lib.MODE_PHARM
        pharm = self.spec.mode in (_lib.MODE_PHARM, _lib.MODE_SHAPE_PHARM)
        mqp, mqh = (int(max_query_points) if shape else 0), (int(max_query_pharm) if pharm else 0)
        # admission: as many query slots as the HBM budget holds, at most max_active
        fixed = arena_bytes(self.launch_pairs, self.B, 0, self.top_k, mqp, mqh)
        slot = arena_bytes(0, 0, 1, self.top_k, mqp, mqh)
        if fixed + slot > hbm_budget_bytes:
            raise MemoryError("the HBM budget does not hold one query slot and one launch")
        self.max_active = max(1, min(int(max_active), (hbm_budget_bytes - fixed) // slot,
                                     self.launch_pairs))
        self.table = restart_table(self.B)
        prm = _lib.FssScreenParams()
        prm.n_restarts, prm.max_steps = self.B, self.opt.max_steps
        prm.lr, prm.beta1, prm.beta2, prm.eps = self.opt.lr, self.opt.beta1, self.opt.beta2, self.opt.eps
        prm.early_stop, prm.es_tol = int(self.opt.early_stop), self.opt.es_tol
        prm.es_patience = int(self.opt.es_patience)
        prm.top_k, prm.launch_pairs, prm.max_active = self.top_k, self.launch_pairs, self.max_active
        prm.max_query_points, prm.max_query_pharm = mqp, mqh
        prm.lanes_per_pose, prm.restarts_per_wg, prm.variant = int(lanes_per_pose), self.B, 0
        prm.seed_flags = _lib.SEED_PHARM if self.spec.mode == _lib.MODE_PHARM else 0
        prm.quat_table = self.table.ctypes.data
        self._sp = self.spec.struct()
        self._h = ctypes.c_void_p()
        _lib.call("fss_screen_create", library.native(), _lib.byref(self._sp), _lib.byref(prm),
                  ctypes.byref(self._h))
        self._futures: Dict[int, Future] = {}
        self._submitted: Dict[int, int] = {}
        self._cv = threading.Condition()
        self._stop = False
        self._poller = threading.Thread(target=self._poll_loop, name="fss-screen-poll", daemon=True)
        self._poller.start()

    # -- public ------------------------------------------------------------------------------
    def submit(self, query: dict, top_k: Optional[int] = None) -> Future:
        """Queue a query (dict with the arrays its mode needs: ``points`` [+ ``esp``] and/or
        ``pharm_types``, ``pharm_anchors``, ``pharm_vectors``).  Resolves to a ScreenResult."""
        k = int(top_k or self.top_k)
        if not 1 <= k <= self.top_k:
            raise ValueError(f"top_k must be in [1, {self.top_k}] for this service")
        q = _lib.FssQuery()
        keep = []
        if self.spec.mode != _lib.MODE_PHARM:
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
            self._cv.notify_all()
        return fut

    @property
    def launches(self) -> int:
        """Launches run so far (the final count once closed)."""
        with self._cv:
            if self._stop:
                return self._final_launches
            n = ctypes.c_int64()
            _lib.call("fss_screen_stats", self._h, ctypes.byref(n), None)
            return int(n.value)

    def close(self) -> None:
        with self._cv:
            if self._stop:
                return
            n = ctypes.c_int64()
            _lib.call("fss_screen_stats", self._h, ctypes.byref(n), None)
            self._final_launches = int(n.value)
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
        while True:
            with self._cv:
                while not self._futures and not self._stop:
                    self._cv.wait()
                if self._stop:
                    return
                tickets = sorted(self._futures)
            # wait (in native code, GIL released) up to 5 ms on the oldest ticket, then sweep
            for i, t in enumerate(tickets):
                rc = lib.fss_screen_poll(self._h, t, 5 if i == 0 else 0, ctypes.byref(hits))
                if rc == _lib.PENDING:
                    continue
                with self._cv:
                    fut = self._futures.pop(t)
                    k = self._submitted.pop(t)
                if rc != 0:
                    fut.set_exception(_lib.FssError("fss_screen_poll", rc, lib.fss_last_error().decode()))
                    continue
                n = min(int(hits.n), k)
                res = [Hit(float(score[i]), self.lib.id_of(int(index[i])), tf[i].reshape(4, 4).copy())
                       for i in range(n)]
                fut.set_result(ScreenResult(res, float(hits.latency_ms), float(hits.queue_ms), self.lib.n))


def _as_library(library, mode: str, device):
    """(ConformerLibrary, owned): a ProfileStore or store path is loaded into HBM first."""
    from .store import ProfileStore, load_library
    if isinstance(library, (ProfileStore, str, os.PathLike)):
        return load_library(library, mode, device=device if device is not None else 0), True
    return library, False


def screen(query: dict, library, mode: str = "vol", *, top_k: int = 1000, device=None,
           **kwargs) -> List[Hit]:
    """One query against the whole library (a ConformerLibrary, a ProfileStore or a store
    path); ``Hit(score, id, transform)`` sorted descending."""
    lib, owned = _as_library(library, mode, device)
    try:
        with ScreenService(lib, mode, top_k=top_k, **kwargs) as svc:
            return svc.submit(query).result().hits
    finally:
        if owned:
            lib.close()


def screen_many(queries: Sequence[dict], library, mode: str = "vol", *, top_k: int = 1000,
                device=None, **kwargs) -> List[List[Hit]]:
    """A panel of queries served together (continuous batching); one hit list per query."""
    lib, owned = _as_library(library, mode, device)
    try:
        with ScreenService(lib, mode, top_k=top_k, **kwargs) as svc:
            futs = [svc.submit(q) for q in queries]
            return [f.result().hits for f in futs]
    finally:
        if owned:
            lib.close()
