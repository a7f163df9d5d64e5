"""Score-as-loss: the hand-written HIP backward vs finite differences (gradcheck, FP64), vs
autograd of the FP64 oracle's scoring restatement, and contracted to SE(3) vs the oracle's
analytic gradient (needs an MI355X)."""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import synth
from fast_shepherd_score_amd.autograd import gaussian_similarity
from fast_shepherd_score_amd.constants import LAM_SCALING
from oracle import fss_oracle as orc

pytestmark = pytest.mark.gpu
LAM = 0.3 * LAM_SCALING


def _clouds(P, N, M, seed, dtype):
    rng = np.random.default_rng(seed)
    ref = rng.normal(size=(P, N, 3)) * 1.5
    fit = rng.normal(size=(P, M, 3)) * 1.5
    re = np.clip(rng.normal(0, 0.1, (P, N)), -0.3, 0.3)
    fe = np.clip(rng.normal(0, 0.1, (P, M)), -0.3, 0.3)
    t = lambda a: torch.tensor(a, dtype=dtype)
    return t(ref), t(fit), t(re), t(fe)


@pytest.mark.parametrize("esp", [False, True])
def test_gradcheck_fp64(dev, esp):
    ref, fit, re, fe = (x.to(dev) for x in _clouds(3, 8, 7, 0, torch.float64))
    ref.requires_grad_(True)
    fit.requires_grad_(True)
    # a small lam makes the ESP weights matter at these charge scales
    lam = 0.05 if esp else LAM
    f = lambda r, x: gaussian_similarity(r, x, re if esp else None, fe if esp else None, 0.81, lam)
    assert torch.autograd.gradcheck(f, (ref, fit), eps=1e-6, atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize("esp", [False, True])
def test_fp32_matches_oracle_autograd(dev, esp):
    P, N, M = 16, 300, 280
    ref64, fit64, re64, fe64 = _clouds(P, N, M, 1, torch.float64)
    lam = 0.5 if esp else LAM
    x = fit64.clone().requires_grad_(True)
    r = ref64.clone().requires_grad_(True)
    if esp:
        S_ref = orc.tanimoto_esp(r, x, re64, fe64, 0.81, lam)
    else:
        S_ref = orc.tanimoto_shape(r, x, 0.81)
    S_ref.sum().backward()
    ref, fit = ref64.float().to(dev).requires_grad_(True), fit64.float().to(dev).requires_grad_(True)
    S = gaussian_similarity(ref, fit, re64.float().to(dev) if esp else None,
                            fe64.float().to(dev) if esp else None, 0.81, lam)
    S.sum().backward()
    np.testing.assert_allclose(S.detach().cpu().numpy(), S_ref.detach().numpy(), atol=1e-5)
    for got, want in ((fit.grad, x.grad), (ref.grad, r.grad)):
        g, w = got.cpu().numpy(), want.numpy()
        np.testing.assert_allclose(g, w, atol=1e-4 * np.abs(w).max())


def test_contraction_to_se3_matches_oracle(dev):
    """Moving the fit rigidly: the coordinate gradient contracted through b' = R b + t equals the
    oracle's SE(3) gradient (after the tangent projection both apply)."""
    rs, rse, fs, fse = synth.surf_esp_pairs(1, atoms=(8, 12), points=(40, 60), seed=4)
    pair = orc.Pair("esp", ref=rs[0], fit=fs[0], ref_esp=rse[0], fit_esp=fse[0], lam=LAM)
    rng = np.random.default_rng(5)
    q = synth.random_quaternions(rng, 1).astype(np.float64)
    q /= np.linalg.norm(q)
    t = rng.normal(0, 0.4, (1, 3))
    se3 = torch.tensor(np.concatenate([q, t], 1))
    _, g_or, _ = pair.loss_grad_se3(se3, pair.self_overlaps())   # d(1-S)/d(q,t), B = 1
    R = orc.rot_from_unit_quat(torch.tensor(q))[0]
    fit0 = torch.tensor(fs[0], dtype=torch.float64)
    moved = (fit0 @ R.T + torch.tensor(t[0])).unsqueeze(0).to(dev).requires_grad_(True)
    S = gaussian_similarity(torch.tensor(rs[0], dtype=torch.float64)[None].to(dev), moved,
                            torch.tensor(rse[0], dtype=torch.float64)[None].to(dev),
                            torch.tensor(fse[0], dtype=torch.float64)[None].to(dev), 0.81, LAM)
    (1 - S).sum().backward()
    gb = moved.grad[0].cpu()                       # dL/db' (M, 3)
    gt = gb.sum(0)
    GR = gb.T @ fit0                               # dL/dR as a general matrix
    qh = torch.tensor(q)
    gq = orc.project_grad_R(GR[None], qh)[0]
    gq = gq - qh[0] * (qh[0] @ gq)                 # tangent part (unit q: |q| = 1)
    np.testing.assert_allclose(gt.numpy(), g_or[0, 4:].numpy(), atol=1e-9)
    np.testing.assert_allclose(gq.numpy(), g_or[0, :4].numpy(), atol=1e-9)


@pytest.mark.parametrize("N,M", [(1, 3), (70, 130), (64, 65)])
def test_ragged_tile_edges_and_determinism(dev, N, M):
    """Cloud sizes that end mid-tile (64-row forward tiles, 16-row backward units split over 4
    column slices) against the FP64 oracle; the backward is bitwise repeatable."""
    ref64, fit64, re64, fe64 = _clouds(3, N, M, 7, torch.float64)
    x = fit64.clone().requires_grad_(True)
    r = ref64.clone().requires_grad_(True)
    orc.tanimoto_esp(r, x, re64, fe64, 0.81, 0.5).sum().backward()
    grads = []
    for _ in range(2):
        ref = ref64.float().to(dev).requires_grad_(True)
        fit = fit64.float().to(dev).requires_grad_(True)
        S = gaussian_similarity(ref, fit, re64.float().to(dev), fe64.float().to(dev), 0.81, 0.5)
        S.sum().backward()
        grads.append((fit.grad.cpu().numpy(), ref.grad.cpu().numpy()))
    S_ref = orc.tanimoto_esp(ref64, fit64, re64, fe64, 0.81, 0.5).numpy()
    np.testing.assert_allclose(S.detach().cpu().numpy(), S_ref, atol=1e-5)
    for got, want in zip(grads[0], (x.grad.numpy(), r.grad.numpy())):
        np.testing.assert_allclose(got, want, atol=1e-4 * np.abs(want).max())
    for a, b in zip(grads[0], grads[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype,N,M", [(torch.float32, 6000, 4500), (torch.float64, 3000, 2500)])
def test_clouds_beyond_lds_stream(dev, dtype, N, M):
    """A pair whose clouds exceed LDS (FP32: 8187 points, FP64: 5114) runs the streamed kernels:
    S and both gradients against autograd of the FP64 oracle (ESP weights on)."""
    print(f"case score-as-loss streamed {dtype} N={N} M={M}", flush=True)
    ref64, fit64, re64, fe64 = _clouds(1, N, M, 5, torch.float64)
    ref64, fit64 = ref64 * 6.0, fit64 * 6.0          # molecule-sized spread for that many points
    x = fit64.clone().requires_grad_(True)
    r = ref64.clone().requires_grad_(True)
    S_ref = orc.tanimoto_esp(r, x, re64, fe64, 0.81, 0.5)
    S_ref.sum().backward()
    ref = ref64.to(dtype).to(dev).requires_grad_(True)
    fit = fit64.to(dtype).to(dev).requires_grad_(True)
    S = gaussian_similarity(ref, fit, re64.to(dtype).to(dev), fe64.to(dtype).to(dev), 0.81, 0.5)
    S.sum().backward()
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    np.testing.assert_allclose(S.detach().cpu().numpy(), S_ref.detach().numpy(), atol=tol)
    gtol = 1e-9 if dtype == torch.float64 else 1e-4
    for got, want in ((fit.grad, x.grad), (ref.grad, r.grad)):
        g, w = got.cpu().numpy(), want.numpy()
        np.testing.assert_allclose(g, w, atol=gtol * np.abs(w).max())


def _score_ragged(clouds, lam, dev, esp=True):
    """fss_score_fwd / fss_score_bwd (FP32; ESP-weighted or not) through the C ABI on a ragged
    batch of (ref (N, 3), fit (M, 3), ref_esp (N,), fit_esp (M,)) pairs: S, overlaps, dS/dfit,
    dS/dref."""
    from fast_shepherd_score_amd import _lib
    pack = lambda xs, ws: torch.cat([torch.cat([x, w[:, None]], 1) for x, w in zip(xs, ws)]).float().to(dev).contiguous()
    rng = lambda ns: torch.tensor([[sum(ns[:i]), n] for i, n in enumerate(ns)], dtype=torch.int32, device=dev)
    rp, fp = pack([c[0] for c in clouds], [c[2] for c in clouds]), pack([c[1] for c in clouds], [c[3] for c in clouds])
    ns, ms = [len(c[0]) for c in clouds], [len(c[1]) for c in clouds]
    rr, fr = rng(ns), rng(ms)
    pb = _lib.FssPairBatch()
    pb.n_pairs, pb.ref_pts, pb.fit_pts, pb.ref_rng, pb.fit_rng = len(clouds), _lib.ptr(rp), _lib.ptr(fp), _lib.ptr(rr), _lib.ptr(fr)
    pb.max_ref, pb.max_fit = max(ns), max(ms)
    sp = _lib.FssScoreParams()
    sp.mode, sp.alpha, sp.lam = (_lib.MODE_SHAPE_ESP if esp else _lib.MODE_SHAPE), 0.81, lam
    P = len(clouds)
    S = torch.empty(P, device=dev)
    ov = torch.empty((P, 4), device=dev)
    stream = int(torch.cuda.current_stream().cuda_stream)
    _lib.call("fss_score_fwd", _lib.byref(pb), _lib.byref(sp), _lib.ptr(S), _lib.ptr(ov), stream)
    g = torch.ones(P, device=dev)
    d_fit = torch.empty((sum(ms), 4), device=dev)
    d_ref = torch.empty((sum(ns), 4), device=dev)
    _lib.call("fss_score_bwd", _lib.byref(pb), _lib.byref(sp), _lib.ptr(ov), _lib.ptr(g), _lib.ptr(d_fit),
              _lib.ptr(d_ref), stream)
    torch.cuda.synchronize()
    return S.cpu(), ov.cpu(), d_fit.cpu(), d_ref.cpu()


@pytest.mark.parametrize("esp", [False, True])
def test_streamed_equals_staged_bitwise(dev, esp):
    """The streamed kernels take the staged kernels' sums in the same order: a small pair gives
    the same bits when its batch also holds a pair beyond LDS (which streams the whole batch)."""
    print(f"case score-as-loss staged vs streamed esp={esp}, 300 x 280 beside 6000 x 4500", flush=True)
    small = [x[0] for x in _clouds(1, 300, 280, 11, torch.float64)]
    big = [x[0] for x in _clouds(1, 6000, 4500, 12, torch.float64)]
    big[0], big[1] = big[0] * 6.0, big[1] * 6.0
    S1, ov1, df1, dr1 = _score_ragged([small], 0.5, dev, esp)
    S2, ov2, df2, dr2 = _score_ragged([small, big], 0.5, dev, esp)
    assert torch.isfinite(S2).all()
    assert torch.equal(S1[0], S2[0]) and torch.equal(ov1[0], ov2[0])
    assert torch.equal(df1, df2[:280]) and torch.equal(dr1, dr2[:300])


@pytest.mark.parametrize("spread,shift", [(4.0, 0.0), (8.0, 0.0), (1.0, 30.0)])
def test_far_points_take_direct_form(dev, spread, shift):
    """Pairs reaching to and beyond the FP32 expanded form's range from the ref centroid (about
    30 A: a wide pair inside it, a wider pair and a displaced fit beyond it, which run the direct
    form): S and both gradients against autograd of the FP64 oracle, staged in LDS (300 x 280
    points)."""
    print(f"case score-as-loss far points spread={spread} shift={shift}", flush=True)
    ref64, fit64, re64, fe64 = _clouds(2, 300, 280, 21, torch.float64)
    ref64, fit64 = ref64 * spread, fit64 * spread
    fit64 = fit64 + torch.tensor([shift, 0.0, 0.0], dtype=torch.float64)
    x = fit64.clone().requires_grad_(True)
    r = ref64.clone().requires_grad_(True)
    S_ref = orc.tanimoto_esp(r, x, re64, fe64, 0.81, 0.5)
    S_ref.sum().backward()
    ref = ref64.float().to(dev).requires_grad_(True)
    fit = fit64.float().to(dev).requires_grad_(True)
    S = gaussian_similarity(ref, fit, re64.float().to(dev), fe64.float().to(dev), 0.81, 0.5)
    S.sum().backward()
    np.testing.assert_allclose(S.detach().cpu().numpy(), S_ref.detach().numpy(), atol=1e-5)
    for got, want in ((fit.grad, x.grad), (ref.grad, r.grad)):
        g, w = got.cpu().numpy(), want.numpy()
        np.testing.assert_allclose(g, w, atol=1e-4 * np.abs(w).max() + 1e-12)
