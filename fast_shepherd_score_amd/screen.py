"""Screening: queries against a conformer library resident in HBM, with continuous batching.

Front-end restating the reference's screening contract (shepherd_score/screen.py:74, 1038-1192):
``screen(query, library, mode, top_k, **align_kwargs) -> [Hit(score, id, transform)]`` sorted by
score, descending.  The reference streams ``ProfileStore`` shards from disk through per-query
aligner calls; here the library is packed once into HBM and a :class:`ScreenService` serves a
stream of queries:

* admission: a request is admitted when an active slot is free and its device workspace fits
  the HBM budget (:class:`Admission`); otherwise it waits in FIFO order;
* continuous batching: every launch is filled from *all* active requests (round-robin slices of
  their remaining library range, :func:`plan_launch`), so a request that arrives while others are
  on the GPU joins the very next launch and its pairs run side by side with theirs;
* per-pair results do not depend on the co-scheduled pairs: the tiling is fixed per service, the
  early stop is per pair, seeds are per pair (device kernel ``fss_seed_poses``), and every
  reduction has a fixed order -- so a query's hits are bit-identical alone and in the mix.
"""
from __future__ import annotations

import collections
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .align import OptimSpec, ScoreSpec, align, self_overlaps
from .constants import LAM_SCALING
from .packing import CloudSet, PairBatch, PharmSet
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
        self.ids = np.arange(n, dtype=np.int64) if ids is None else np.asarray(ids, np.int64)

    @classmethod
    def from_arrays(cls, points=None, esp=None, pharm_types=None, pharm_anchors=None,
                    pharm_vectors=None, ids=None, device="cuda") -> "ConformerLibrary":
        shape = CloudSet.from_arrays(points, esp, device=device) if points is not None else None
        pharm = (PharmSet.from_arrays(pharm_types, pharm_anchors, pharm_vectors, device=device)
                 if pharm_types is not None else None)
        return cls(shape, pharm, ids)

    def nbytes(self) -> int:
        tot = 0
        for cs in (self.shape, self.pharm):
            if cs is None:
                continue
            for t in vars(cs).values():
                if isinstance(t, torch.Tensor):
                    tot += t.numel() * t.element_size()
        return tot


class Admission:
    """HBM budget accounting for active requests (pure host logic)."""

    def __init__(self, budget_bytes: int, max_active: int):
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.used = 0
        self.active = 0

    def try_admit(self, nbytes: int) -> bool:
        if self.active >= self.max_active or self.used + nbytes > self.budget:
            return False
        self.used += nbytes
        self.active += 1
        return True

    def release(self, nbytes: int) -> None:
        self.used -= nbytes
        self.active -= 1


@dataclass
class _Request:
    rid: int
    query: dict
    top_k: int
    t_submit: float
    future: Future
    nbytes: int = 0
    next: int = 0                    # next library index to schedule
    done: int = 0                    # library entries finished
    best_s: Optional[torch.Tensor] = None
    best_i: Optional[torch.Tensor] = None
    best_tf: Optional[torch.Tensor] = None
    t_start: float = 0.0


def plan_launch(remaining: Sequence[int], capacity: int) -> List[int]:
    """Split ``capacity`` library entries over requests with ``remaining`` entries each:
    water-filling, so every active request advances in every launch (continuous batching)."""
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


@dataclass
class ScreenResult:
    hits: List[Hit]
    latency_ms: float
    queue_ms: float
    pairs: int


class ScreenService:
    """Serve screening queries against a resident library with continuous batching."""

    def __init__(self, library: ConformerLibrary, mode: str = "vol", *, num_repeats: int = 10,
                 max_num_steps: int = 100, lr: float = 0.1, top_k: int = 100,
                 launch_pairs: int = 65536, hbm_budget_bytes: int = 8 << 30, max_active: int = 16,
                 early_stop: bool = True, alpha: float = 0.81, lam: Optional[float] = None,
                 similarity: str = "tanimoto", directionless: bool = True,
                 lanes_per_pose: int = 4, device=None):
        if mode not in _KIND:
            raise ValueError(f"unsupported screening mode {mode!r}")
        self.lib, self.mode = library, mode
        self.dev = torch.device(device) if device is not None else (
            library.shape.pts.device if library.shape is not None else library.pharm.anc.device)
        if lam is None:
            lam = 0.3 * LAM_SCALING if mode == "surf_esp" else 0.1
        self.spec = ScoreSpec(mode=_KIND[mode], alpha=alpha, lam=lam, similarity=similarity,
                              directionless=directionless)
        self.B = int(num_repeats)
        self.opt = OptimSpec(max_steps=int(max_num_steps), lr=float(lr), early_stop=bool(early_stop))
        self.tiling = dict(lanes_per_pose=int(lanes_per_pose), restarts_per_wg=self.B)
        self.top_k = int(top_k)
        self.launch_pairs = int(launch_pairs)
        self.admission = Admission(hbm_budget_bytes, max_active)
        self.table = torch.from_numpy(restart_table(self.B)).to(self.dev)
        self.stream = torch.cuda.Stream(device=self.dev)
        self._inbox: "queue.Queue[_Request]" = queue.Queue()
        self._waiting: collections.deque = collections.deque()
        self._active: List[_Request] = []
        self._next_id = 0
        self._stop = threading.Event()
        self._lock = threading.Lock()
        self.launches = 0
        self._thread = threading.Thread(target=self._loop, name="fss-screen", daemon=True)
        self._thread.start()

    # -- public ------------------------------------------------------------------------------
    def submit(self, query: dict, top_k: Optional[int] = None) -> Future:
        """Queue a query (dict with the arrays its mode needs: ``points`` [+ ``esp``] and/or
        ``pharm_types``, ``pharm_anchors``, ``pharm_vectors``).  Resolves to a ScreenResult."""
        fut: Future = Future()
        with self._lock:
            rid = self._next_id
            self._next_id += 1
        self._inbox.put(_Request(rid, query, int(top_k or self.top_k), time.perf_counter(), fut))
        return fut

    def close(self) -> None:
        self._stop.set()
        self._inbox.put(None)
        self._thread.join(timeout=60)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- scheduler ----------------------------------------------------------------------------
    def _workspace_bytes(self, req: _Request) -> int:
        k = min(req.top_k, self.lib.n)
        npts = sum(len(np.asarray(v)) for key, v in req.query.items() if key in ("points", "pharm_anchors"))
        return 64 * npts + k * (8 + 8 + 64) + 4096

    def _drain_inbox(self, block: bool) -> None:
        try:
            item = self._inbox.get(timeout=0.05) if block else self._inbox.get_nowait()
            while True:
                if item is not None:
                    self._waiting.append(item)
                item = self._inbox.get_nowait()
        except queue.Empty:
            pass

    def _admit(self) -> None:
        while self._waiting:
            req = self._waiting[0]
            req.nbytes = self._workspace_bytes(req)
            if req.nbytes > self.admission.budget:
                self._waiting.popleft()
                req.future.set_exception(MemoryError("query workspace exceeds the HBM budget"))
                continue
            if not self.admission.try_admit(req.nbytes):
                break
            self._waiting.popleft()
            req.t_start = time.perf_counter()
            self._active.append(req)

    def _loop(self) -> None:
        torch.cuda.set_device(self.dev)
        while not self._stop.is_set():
            self._drain_inbox(block=not self._active)
            self._admit()
            if not self._active:
                continue
            try:
                self._launch()
            except Exception as exc:   # fail the requests of this launch, keep serving
                for r in self._active:
                    if not r.future.done():
                        r.future.set_exception(exc)
                    self.admission.release(r.nbytes)
                self._active = []

    def _query_sets(self, reqs: Sequence[_Request]):
        qs = [r.query for r in reqs]
        shape = pharm = None
        if self.lib.shape is not None:
            shape = CloudSet.from_arrays([q["points"] for q in qs],
                                         [q.get("esp") for q in qs] if self.spec.mode == _lib.MODE_SHAPE_ESP else None,
                                         device=self.dev)
        if self.lib.pharm is not None:
            pharm = PharmSet.from_arrays([q["pharm_types"] for q in qs], [q["pharm_anchors"] for q in qs],
                                         [q["pharm_vectors"] for q in qs], device=self.dev)
        return shape, pharm

    def _launch(self) -> None:
        reqs = list(self._active)
        take = plan_launch([self.lib.n - r.next for r in reqs], self.launch_pairs)
        reqs = [r for r, k in zip(reqs, take) if k > 0]
        take = [k for k in take if k > 0]
        ref_idx = np.concatenate([np.full(k, i, np.int64) for i, k in enumerate(take)])
        fit_idx = np.concatenate([np.arange(r.next, r.next + k) for r, k in zip(reqs, take)])
        with torch.cuda.stream(self.stream):
            qshape, qpharm = self._query_sets(reqs)
            pb = PairBatch(qshape, self.lib.shape, qpharm, self.lib.pharm, ref_idx=ref_idx,
                           fit_idx=fit_idx)
            P = pb.n_pairs
            q0 = torch.empty((P, self.B, 4), dtype=torch.float32, device=self.dev)
            t0 = torch.empty((P, self.B, 3), dtype=torch.float32, device=self.dev)
            bs = pb.struct()
            _lib.call("fss_seed_poses", _lib.byref(bs), int(self.spec.mode == _lib.MODE_PHARM), self.B,
                      _lib.ptr(self.table), _lib.ptr(q0), _lib.ptr(t0), int(self.stream.cuda_stream))
            so = self_overlaps(pb, self.spec, stream=self.stream)
            res = align(pb, self.spec, q0, t0, self.opt, order="none", self_ov=so,
                        stream=self.stream, **self.tiling)
            off = 0
            for r, k in zip(reqs, take):
                s = res.score[off:off + k]
                idx = torch.arange(r.next, r.next + k, device=self.dev)
                tf = res.transform[off:off + k]
                if r.best_s is not None:
                    s = torch.cat([r.best_s, s])
                    idx = torch.cat([r.best_i, idx])
                    tf = torch.cat([r.best_tf, tf])
                kk = min(r.top_k, s.shape[0])
                top = torch.topk(s, kk, sorted=True)
                r.best_s, r.best_i, r.best_tf = top.values, idx[top.indices], tf[top.indices]
                off += k
            done_evt = torch.cuda.Event()
            done_evt.record(self.stream)
        done_evt.synchronize()
        self.launches += 1
        now = time.perf_counter()
        finished = []
        for r, k in zip(reqs, take):
            r.next += k
            r.done += k
            if r.done >= self.lib.n:
                finished.append(r)
        for r in finished:
            sc = r.best_s.cpu().numpy()
            ii = r.best_i.cpu().numpy()
            tf = r.best_tf.cpu().numpy()
            hits = [Hit(float(sc[i]), int(self.lib.ids[ii[i]]), tf[i]) for i in range(len(sc))]
            r.future.set_result(ScreenResult(hits, (now - r.t_submit) * 1e3,
                                             (r.t_start - r.t_submit) * 1e3, self.lib.n))
            self._active.remove(r)
            self.admission.release(r.nbytes)


def screen(query: dict, library: ConformerLibrary, mode: str = "vol", *, top_k: int = 1000,
           **kwargs) -> List[Hit]:
    """One query against the whole library; ``Hit(score, id, transform)`` sorted descending."""
    with ScreenService(library, mode, top_k=top_k, **kwargs) as svc:
        return svc.submit(query).result().hits


def screen_many(queries: Sequence[dict], library: ConformerLibrary, mode: str = "vol", *,
                top_k: int = 1000, **kwargs) -> List[List[Hit]]:
    """A panel of queries served together (continuous batching); one hit list per query."""
    with ScreenService(library, mode, top_k=top_k, **kwargs) as svc:
        futs = [svc.submit(q) for q in queries]
        return [f.result().hits for f in futs]
