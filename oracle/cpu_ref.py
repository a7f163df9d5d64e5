"""ctypes wrapper of oracle/libfss_cpu.so, the C++ (std::thread pool, vectorised kernel) CPU baseline
(test infrastructure)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from fast_shepherd_score_amd import _lib as _abi   # ctypes layouts of include/fss.h only

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfss_cpu.so")
_cdll = None


def load():
    global _cdll
    if _cdll is None:
        if not os.path.exists(LIB):
            from oracle.build import build_oracle
            build_oracle()
        _cdll = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        _cdll.fss_cpu_self_overlap.argtypes = [P, P, P, ctypes.c_int]
        _cdll.fss_cpu_align.argtypes = [P, P, P, P, P, P, P, P, P, P, ctypes.c_int]
    return _cdll


def _pack(clouds, w=None):
    counts = np.array([len(c) for c in clouds], np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]) if len(counts) else counts
    pts = np.zeros((max(int(counts.sum()), 1), 4), np.float32)
    for i, c in enumerate(clouds):
        pts[starts[i]:starts[i] + counts[i], :3] = np.asarray(c, np.float32).reshape(-1, 3)
        if w is not None:
            pts[starts[i]:starts[i] + counts[i], 3] = np.asarray(w[i], np.float32)
    return pts, np.stack([starts, counts], 1).astype(np.int32)


class HostBatch:
    """Host-memory packed pairs for the CPU baseline (same layout as the device PairBatch)."""

    def __init__(self, ref=None, fit=None, ref_w=None, fit_w=None, ph=None):
        self.keep = []
        s = _abi.FssPairBatch()
        n = len(fit) if fit is not None else len(ph["fit_types"])
        s.n_pairs = n
        if ref is not None:
            rp, rr = _pack(ref, ref_w)
            fp, fr = _pack(fit, fit_w)
            self.keep += [rp, rr, fp, fr]
            s.ref_pts, s.ref_rng, s.fit_pts, s.fit_rng = (a.ctypes.data for a in (rp, rr, fp, fr))
            s.max_ref, s.max_fit = int(rr[:, 1].max()), int(fr[:, 1].max())
        if ph is not None:
            for side in ("ref", "fit"):
                anc, rng = _pack(ph[f"{side}_anchors"])
                vec, _ = _pack(ph[f"{side}_vectors"])
                typ = np.zeros(anc.shape[0], np.int32)
                for i, t in enumerate(ph[f"{side}_types"]):
                    typ[rng[i, 0]:rng[i, 0] + rng[i, 1]] = t
                self.keep += [anc, vec, typ, rng]
                setattr(s, f"{side}_anc", anc.ctypes.data)
                setattr(s, f"{side}_vec", vec.ctypes.data)
                setattr(s, f"{side}_typ", typ.ctypes.data)
                setattr(s, f"{side}_prng", rng.ctypes.data)
        self.s = s
        self.n = n


def align(hb: HostBatch, spec, q0: np.ndarray, t0: np.ndarray, opt, n_threads: int):
    """Run the CPU baseline; q0 (n, B, 4), t0 (n, B, 3).  Returns dict of numpy outputs."""
    lib = load()
    n, B = q0.shape[:2]
    sp = spec.struct()
    ap = _abi.FssAlignParams()
    ap.n_restarts, ap.max_steps = B, opt.max_steps
    ap.lr, ap.beta1, ap.beta2, ap.eps = opt.lr, opt.beta1, opt.beta2, opt.eps
    ap.early_stop, ap.es_tol, ap.es_patience = int(opt.early_stop), opt.es_tol, opt.es_patience
    so = np.zeros((n, 4), np.float64)
    lib.fss_cpu_self_overlap(ctypes.byref(hb.s), ctypes.byref(sp), so.ctypes.data, n_threads)
    q = np.ascontiguousarray(q0, np.float32).copy()
    t = np.ascontiguousarray(t0, np.float32).copy()
    ps = np.zeros((n, B), np.float64)
    best = np.zeros(n, np.float64)
    bidx = np.zeros(n, np.int32)
    nst = np.zeros(n, np.int32)
    lib.fss_cpu_align(ctypes.byref(hb.s), ctypes.byref(sp), ctypes.byref(ap), so.ctypes.data,
                      q.ctypes.data, t.ctypes.data, ps.ctypes.data, best.ctypes.data,
                      bidx.ctypes.data, nst.ctypes.data, n_threads)
    return {"score": best, "best_idx": bidx, "pose_score": ps, "q": q, "t": t, "n_steps": nst,
            "self_ov": so}
