"""The per-pair early-stop rule on the GPU, pinned by closed-form step counts (needs an MI355X).

Each case keeps every restart's loss exactly constant with an exactly zero gradient, so the rule
alone fixes how many steps run (tests/_cases.py: 12 from the optimum, 11 from L = 1 with the
analytic optimisers' start value 1.0, 12 for the ShaEP optimiser's +inf start).  The persistent
kernel (every inner-loop variant), the stepped path (eager and graph-replayed) and the plain
per-pair path must all report the oracle's count.
"""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import _lib
from fast_shepherd_score_amd.align import OptimSpec, ScoreSpec, SteppedAligner, align, align_plain
from fast_shepherd_score_amd.constants import LAM_SCALING
from fast_shepherd_score_amd.packing import CloudSet, PairBatch, ShaepSet
from tests._cases import ES_EXPECTED, es_seeds, es_shape_clouds, es_shaep_molecules, to_dev

pytestmark = pytest.mark.gpu
B = 3


def _batch(name, dev):
    if name == "shaep_apart":
        a, b = es_shaep_molecules()
        se = lambda m: ShaepSet.from_arrays([m["surf_pos"]], [m["surf_esp"]], [m["centers_w_H"]],
                                            [m["partial_charges"]], [m["radii"]], device=dev)
        pb = PairBatch(CloudSet.from_arrays([a["atom_pos"]], device=dev),
                       CloudSet.from_arrays([b["atom_pos"]], device=dev), ref_se=se(a), fit_se=se(b))
        return pb, ScoreSpec(mode=_lib.MODE_ESP_COMBO, alpha=0.81, lam=LAM_SCALING * 0.001)
    ref, fit = es_shape_clouds(name)
    pb = PairBatch(CloudSet.from_arrays([ref.astype(np.float32)], device=dev),
                   CloudSet.from_arrays([fit.astype(np.float32)], device=dev))
    return pb, ScoreSpec(mode=_lib.MODE_SHAPE)


@pytest.mark.parametrize("name", ["optimum", "apart", "shaep_apart"])
def test_early_stop_step_counts(dev, name):
    print(f"case early stop {name}", flush=True)
    pb, spec = _batch(name, dev)
    q, t = es_seeds(name, B)
    qd, td = to_dev(q, dev), to_dev(t, dev)
    opt = OptimSpec(max_steps=50, early_stop=True)
    want = ES_EXPECTED[name]
    variants = (2,) if spec.mode == _lib.MODE_ESP_COMBO else (1, 2, 3)
    for v in variants:
        r = align(pb, spec, qd, td, opt, variant=v)
        assert int(r.n_steps[0]) == want, ("persistent", v, int(r.n_steps[0]))
        assert torch.equal(r.q, qd) and torch.equal(r.t, td), v     # no pose moved
    for graph in (False, True):
        r = SteppedAligner(pb, spec, qd, td, opt).run(use_graph=graph)
        assert int(r.n_steps[0]) == want, ("stepped", graph, int(r.n_steps[0]))
    r = align_plain(pb, spec, qd, td, opt)
    assert int(r.n_steps[0]) == want, ("plain", int(r.n_steps[0]))
