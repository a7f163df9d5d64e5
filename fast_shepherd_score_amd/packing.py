"""Packed, device-resident cloud layout consumed by libfss_hip.so.

Every cloud of a batch is a run of float4 rows (x, y, z, w) in one contiguous device tensor
(w = ESP value for surfaces / charges, 0 otherwise); a (start, count) row per cloud names it.
Pairs point at a ref cloud and a fit cloud by index, so one query cloud can serve a whole
library slice without copies.  This replaces the reference's padded (K, N_pad, 3) buffers and
N_real / M_real masks (accel/batch/_pad.py:100-125, accel/batch/aligners.py:45-80).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib


def _as_np(a, dtype=np.float32):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


@dataclass
class CloudSet:
    """A packed set of point clouds on one device."""
    pts: torch.Tensor          # (rows, 4) float32
    rng: torch.Tensor          # (n, 2) int32 device: start row, count
    counts: np.ndarray         # (n,) host copy of the counts

    @property
    def n(self) -> int:
        return int(self.counts.shape[0])

    @classmethod
    def from_arrays(cls, xyz: Sequence, w: Optional[Sequence] = None, device="cuda") -> "CloudSet":
        counts = np.array([len(x) for x in xyz], dtype=np.int64)
        starts = np.zeros_like(counts)
        if len(counts) > 1:
            starts[1:] = np.cumsum(counts)[:-1]
        rows = int(counts.sum())
        host = np.zeros((max(rows, 1), 4), dtype=np.float32)
        for i, x in enumerate(xyz):
            if counts[i] == 0:
                continue
            x = _as_np(x).reshape(-1, 3)
            host[starts[i]:starts[i] + counts[i], :3] = x
            if w is not None and w[i] is not None:
                host[starts[i]:starts[i] + counts[i], 3] = _as_np(w[i]).reshape(-1)
        rng = np.stack([starts, counts], axis=1).astype(np.int32)
        return cls(torch.from_numpy(host).to(device), torch.from_numpy(rng).to(device),
                   counts.astype(np.int64))

    @classmethod
    def from_flat(cls, xyz: np.ndarray, counts: np.ndarray, w: Optional[np.ndarray] = None,
                  device="cuda") -> "CloudSet":
        """From concatenated rows (rows, 3) and per-cloud counts (no per-cloud Python loop)."""
        counts = np.asarray(counts, dtype=np.int64)
        rows = int(counts.sum())
        host = np.zeros((max(rows, 1), 4), dtype=np.float32)
        host[:rows, :3] = np.asarray(xyz, np.float32).reshape(-1, 3)
        if w is not None:
            host[:rows, 3] = np.asarray(w, np.float32).reshape(-1)
        starts = np.zeros_like(counts)
        if len(counts) > 1:
            starts[1:] = np.cumsum(counts)[:-1]
        rng = np.stack([starts, counts], axis=1).astype(np.int32)
        return cls(torch.from_numpy(host).to(device), torch.from_numpy(rng).to(device), counts)

    @classmethod
    def from_padded(cls, xyz: np.ndarray, counts: np.ndarray, w: Optional[np.ndarray] = None,
                    device="cuda") -> "CloudSet":
        """From a padded (n, P, 3) array and real counts (vectorised packing)."""
        n, P = xyz.shape[:2]
        counts = np.asarray(counts, dtype=np.int64)
        mask = np.arange(P)[None, :] < counts[:, None]
        host = np.zeros((max(int(counts.sum()), 1), 4), dtype=np.float32)
        host[:int(counts.sum()), :3] = xyz[mask]
        if w is not None:
            host[:int(counts.sum()), 3] = w[mask]
        starts = np.zeros_like(counts)
        if n > 1:
            starts[1:] = np.cumsum(counts)[:-1]
        rng = np.stack([starts, counts], axis=1).astype(np.int32)
        return cls(torch.from_numpy(host).to(device), torch.from_numpy(rng).to(device), counts)


@dataclass
class PharmSet:
    """A packed set of pharmacophore clouds: anchors, direction vectors and int32 types."""
    anc: torch.Tensor          # (rows, 4) float32
    vec: torch.Tensor          # (rows, 4) float32
    typ: torch.Tensor          # (rows,) int32
    rng: torch.Tensor          # (n, 2) int32
    counts: np.ndarray

    @property
    def n(self) -> int:
        return int(self.counts.shape[0])

    @classmethod
    def from_arrays(cls, types: Sequence, anchors: Sequence, vectors: Sequence,
                    device="cuda") -> "PharmSet":
        counts = np.array([len(t) for t in types], dtype=np.int64)
        starts = np.zeros_like(counts)
        if len(counts) > 1:
            starts[1:] = np.cumsum(counts)[:-1]
        rows = max(int(counts.sum()), 1)
        anc = np.zeros((rows, 4), np.float32)
        vec = np.zeros((rows, 4), np.float32)
        typ = np.zeros((rows,), np.int32)
        for i in range(len(counts)):
            if counts[i] == 0:
                continue
            s, c = starts[i], counts[i]
            typ[s:s + c] = _as_np(types[i], np.int64).reshape(-1).astype(np.int32)
            anc[s:s + c, :3] = _as_np(anchors[i]).reshape(-1, 3)
            vec[s:s + c, :3] = _as_np(vectors[i]).reshape(-1, 3)
        rng = np.stack([starts, counts], axis=1).astype(np.int32)
        dev = torch.device(device)
        return cls(torch.from_numpy(anc).to(dev), torch.from_numpy(vec).to(dev),
                   torch.from_numpy(typ).to(dev), torch.from_numpy(rng).to(dev), counts)


@dataclass
class ShaepSet:
    """Packed ShaEP channel of a set of molecules: surface observer points with their ESP
    (x, y, z, esp) and atoms including hydrogens with partial charges (x, y, z, q) and vdW radii."""
    srf: CloudSet              # surface rows, w = ESP at the point
    atm: CloudSet              # atom rows, w = partial charge
    rad: torch.Tensor          # (atom rows,) float32 vdW radii

    @property
    def n(self) -> int:
        return self.srf.n

    @classmethod
    def from_arrays(cls, surf: Sequence, surf_esp: Sequence, atoms: Sequence, charges: Sequence,
                    radii: Sequence, device="cuda") -> "ShaepSet":
        srf = CloudSet.from_arrays(surf, surf_esp, device=device)
        atm = CloudSet.from_arrays(atoms, charges, device=device)
        rows = max(int(atm.counts.sum()), 1)
        r = np.zeros(rows, np.float32)
        if int(atm.counts.sum()):
            r[:int(atm.counts.sum())] = np.concatenate([_as_np(x).reshape(-1) for x in radii])
        return cls(srf, atm, torch.from_numpy(r).to(device))


class PairBatch:
    """Pairs (ref cloud i, fit cloud j) over packed clouds, ready for the C ABI.

    ``ref_idx`` / ``fit_idx`` give, per pair, the cloud index in the ref / fit sets (default:
    pair p uses cloud p of each set).  Shape and pharmacophore channels are optional; the mode
    decides which must be present.
    """

    def __init__(self, ref: Optional[CloudSet] = None, fit: Optional[CloudSet] = None,
                 ref_ph: Optional[PharmSet] = None, fit_ph: Optional[PharmSet] = None,
                 ref_idx=None, fit_idx=None, n_pairs: Optional[int] = None,
                 ref_se: Optional[ShaepSet] = None, fit_se: Optional[ShaepSet] = None):
        self.ref, self.fit, self.ref_ph, self.fit_ph = ref, fit, ref_ph, fit_ph
        self.ref_se, self.fit_se = ref_se, fit_se
        some = fit if fit is not None else fit_ph
        if some is None:
            raise ValueError("a PairBatch needs a fit cloud set")
        if n_pairs is None:
            n_pairs = len(fit_idx) if fit_idx is not None else some.n
        self.n_pairs = int(n_pairs)
        self.ref_idx = None if ref_idx is None else np.asarray(ref_idx, dtype=np.int64)
        self.fit_idx = None if fit_idx is None else np.asarray(fit_idx, dtype=np.int64)
        dev = some.rng.device
        self.device = dev

        def pair_rows(cs, idx):
            if cs is None:
                return None, None
            if idx is None:
                if cs.n != self.n_pairs:
                    raise ValueError(f"cloud set has {cs.n} clouds for {self.n_pairs} pairs")
                return cs.rng, cs.counts
            it = torch.from_numpy(idx).to(dev)
            return cs.rng.index_select(0, it).contiguous(), cs.counts[idx]

        self.ref_rng, self.ref_counts = pair_rows(ref, self.ref_idx)
        self.fit_rng, self.fit_counts = pair_rows(fit, self.fit_idx)
        self.ref_prng, self.ref_pcounts = pair_rows(ref_ph, self.ref_idx)
        self.fit_prng, self.fit_pcounts = pair_rows(fit_ph, self.fit_idx)
        se = lambda x, k: None if x is None else getattr(x, k)
        self.ref_srng, self.ref_scounts = pair_rows(se(ref_se, "srf"), self.ref_idx)
        self.fit_srng, self.fit_scounts = pair_rows(se(fit_se, "srf"), self.fit_idx)
        self.ref_arng, self.ref_acounts = pair_rows(se(ref_se, "atm"), self.ref_idx)
        self.fit_arng, self.fit_acounts = pair_rows(se(fit_se, "atm"), self.fit_idx)

    def _max(self, c):
        return int(c.max()) if c is not None and len(c) else 0

    def struct(self) -> _lib.FssPairBatch:
        s = _lib.FssPairBatch()
        s.n_pairs = self.n_pairs
        p = _lib.ptr
        if self.ref is not None:
            s.ref_pts, s.ref_rng = p(self.ref.pts), p(self.ref_rng)
        if self.fit is not None:
            s.fit_pts, s.fit_rng = p(self.fit.pts), p(self.fit_rng)
        if self.ref_ph is not None:
            s.ref_anc, s.ref_vec = p(self.ref_ph.anc), p(self.ref_ph.vec)
            s.ref_typ, s.ref_prng = p(self.ref_ph.typ), p(self.ref_prng)
        if self.fit_ph is not None:
            s.fit_anc, s.fit_vec = p(self.fit_ph.anc), p(self.fit_ph.vec)
            s.fit_typ, s.fit_prng = p(self.fit_ph.typ), p(self.fit_prng)
        s.max_ref, s.max_fit = self._max(self.ref_counts), self._max(self.fit_counts)
        s.max_ref_ph, s.max_fit_ph = self._max(self.ref_pcounts), self._max(self.fit_pcounts)
        if self.ref_se is not None:
            s.ref_srf, s.ref_srng = p(self.ref_se.srf.pts), p(self.ref_srng)
            s.ref_atm, s.ref_rad, s.ref_arng = p(self.ref_se.atm.pts), p(self.ref_se.rad), p(self.ref_arng)
        if self.fit_se is not None:
            s.fit_srf, s.fit_srng = p(self.fit_se.srf.pts), p(self.fit_srng)
            s.fit_atm, s.fit_rad, s.fit_arng = p(self.fit_se.atm.pts), p(self.fit_se.rad), p(self.fit_arng)
        s.max_ref_srf, s.max_fit_srf = self._max(self.ref_scounts), self._max(self.fit_scounts)
        s.max_ref_atm, s.max_fit_atm = self._max(self.ref_acounts), self._max(self.fit_acounts)
        return s

    def cost(self, mode: int) -> np.ndarray:
        """Pair-terms per pose of every pair (the work the kernel does per step)."""
        c = np.zeros(self.n_pairs, dtype=np.float64)
        if mode != _lib.MODE_PHARM and self.ref_counts is not None:
            c += self.ref_counts.astype(np.float64) * self.fit_counts
        if mode in (_lib.MODE_PHARM, _lib.MODE_SHAPE_PHARM) and self.ref_pcounts is not None:
            c += self.ref_pcounts.astype(np.float64) * self.fit_pcounts
        if mode == _lib.MODE_ESP_COMBO and self.ref_scounts is not None:   # two passes, two directions
            c += 2.0 * (self.ref_scounts.astype(np.float64) * self.fit_acounts
                        + self.fit_scounts.astype(np.float64) * self.ref_acounts)
        return c
