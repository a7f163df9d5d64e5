"""The C++ CPU port (oracle/cpu_baseline.cpp, the bench's cpu_baseline leg) against the FP64
oracle: self-overlaps and best-restart scores from the same seeds, in every objective mode
including pharmacophore extended points.  CPU only."""
import numpy as np
import pytest

from fast_shepherd_score_amd import _lib, synth
from fast_shepherd_score_amd.align import OptimSpec, ScoreSpec
from fast_shepherd_score_amd.constants import LAM_SCALING
from oracle import cpu_ref
from oracle import fss_oracle as orc
from tests._cases import random_seeds

CASES = {
    "shape": dict(mode=_lib.MODE_SHAPE),
    "esp": dict(mode=_lib.MODE_SHAPE_ESP),
    "pharm": dict(mode=_lib.MODE_PHARM),
    "pharm_ext": dict(mode=_lib.MODE_PHARM, extended_points=True),
    "pharm_only_ext": dict(mode=_lib.MODE_PHARM, extended_points=True, only_extended=True,
                           similarity="tversky"),
    "shape_pharm_ext": dict(mode=_lib.MODE_SHAPE_PHARM, extended_points=True),
}
NAME = {_lib.MODE_SHAPE: "shape", _lib.MODE_SHAPE_ESP: "esp", _lib.MODE_PHARM: "pharm",
        _lib.MODE_SHAPE_PHARM: "shape_pharm"}


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


@pytest.mark.parametrize("case", list(CASES))
def test_cpu_port_matches_oracle(case):
    kw = dict(CASES[case])
    mode = kw.pop("mode")
    n, B, steps = 3, 6, 50
    lam = 0.3 * LAM_SCALING
    spec = ScoreSpec(mode=mode, lam=lam, **kw)
    if mode == _lib.MODE_SHAPE:
        ref, fit = synth.vol_pairs(n, 12, 20, seed=3)
        hb = cpu_ref.HostBatch(ref=ref, fit=fit)
        pairs = [orc.Pair("shape", ref=_f32(r), fit=_f32(f)) for r, f in zip(ref, fit)]
    elif mode == _lib.MODE_SHAPE_ESP:
        r, re, f, fe = synth.surf_esp_pairs(n, atoms=(6, 10), points=(20, 40), seed=4)
        hb = cpu_ref.HostBatch(ref=r, fit=f, ref_w=re, fit_w=fe)
        pairs = [orc.Pair("esp", ref=_f32(a), fit=_f32(b), ref_esp=_f32(c), fit_esp=_f32(d), lam=lam)
                 for a, c, b, d in zip(r, re, f, fe)]
    else:
        d = synth.pharm_shape_pairs(n, atoms=(10, 20), pharms=(5, 12), seed=5)
        ph = {k: d[k] for k in ("ref_types", "fit_types", "ref_anchors", "fit_anchors", "ref_vectors",
                                "fit_vectors")}
        shape = mode == _lib.MODE_SHAPE_PHARM
        hb = cpu_ref.HostBatch(ref=d["ref_atoms"] if shape else None, fit=d["fit_atoms"] if shape else None,
                               ph=ph)
        pairs = [orc.Pair(NAME[mode], ref=_f32(d["ref_atoms"][i]), fit=_f32(d["fit_atoms"][i]),
                          ref_types=d["ref_types"][i], fit_types=d["fit_types"][i],
                          ref_anc=_f32(d["ref_anchors"][i]), fit_anc=_f32(d["fit_anchors"][i]),
                          ref_vec=_f32(d["ref_vectors"][i]), fit_vec=_f32(d["fit_vectors"][i]),
                          similarity=kw.get("similarity", "tanimoto"),
                          extended_points=kw.get("extended_points", False),
                          only_extended=kw.get("only_extended", False)) for i in range(n)]
    q, t = random_seeds(n, B, seed=6, t_scale=0.3)
    res = cpu_ref.align(hb, spec, q, t, OptimSpec(max_steps=steps, lr=0.1), n_threads=2)
    for i, p in enumerate(pairs):
        so = p.self_overlaps()
        col = [("VAA_s", 0), ("VBB_s", 1), ("VAA_p", 2), ("VBB_p", 3)]
        for key, c in col:
            if key in so:
                assert res["self_ov"][i, c] == pytest.approx(float(so[key]), rel=1e-6), (key, i)
        o = orc.optimize(p, q[i], t[i], lr=0.1, max_steps=steps)
        assert abs(res["score"][i] - o["score"]) < 1e-3, (case, i, res["score"][i], o["score"])
