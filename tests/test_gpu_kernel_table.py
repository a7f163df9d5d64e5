"""The reference's accelerated kernel table (fast_shepherd_score_amd.kernels) on the GPU against
the FP64 oracle (needs an MI355X).  Padded rows past the real counts hold nearby coordinates,
so a kernel that read them would fail.  Tolerances: values 2e-5 relative, gradients 1e-4 of
the largest component (FP32 kernels); the Adam step 1e-6 against a float32 torch restatement
of the reference's own update (cpu.py:112-138)."""
import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import kernels as kt
from fast_shepherd_score_amd import synth
from fast_shepherd_score_amd.constants import LAM_SCALING
from oracle import fss_oracle as orc

pytestmark = pytest.mark.gpu
DT = torch.float64


def _padded(rng, K, N, real):
    """(K, N, 3) float32 clouds whose rows past real[k] are near-copies of the real ones."""
    X = np.zeros((K, N, 3), np.float32)
    for k in range(K):
        X[k, :real[k]] = synth.random_walks(rng, [int(real[k])])[0]
        X[k, real[k]:] = X[k, :N - real[k]] + rng.normal(0, 0.3, size=(N - real[k], 3))
    return X


def _poses(rng, K):
    q = synth.random_quaternions(rng, K).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.normal(0, 0.6, size=(K, 3)).astype(np.float32)
    return q, t


def _check(V, dQ, dT, O, gQ, gT):
    np.testing.assert_allclose(V, O, rtol=2e-5, atol=1e-6)
    scale = max(np.abs(gQ).max(), np.abs(gT).max(), 1e-6)
    np.testing.assert_allclose(dQ, gQ, atol=1e-4 * scale)
    np.testing.assert_allclose(dT, gT, atol=1e-4 * scale)


@pytest.mark.parametrize("esp", [False, True])
def test_shape_and_esp_overlap_grad(dev, esp):
    print(f"case kernel table shape esp={esp}", flush=True)
    rng = np.random.default_rng(1)
    K, N, M = 5, 30, 26
    Nr, Mr = rng.integers(8, N + 1, size=K), rng.integers(8, M + 1, size=K)
    A, B = _padded(rng, K, N, Nr), _padded(rng, K, M, Mr)
    cA = rng.normal(0, 0.2, size=(K, N)).astype(np.float32)
    cB = rng.normal(0, 0.2, size=(K, M)).astype(np.float32)
    q, t = _poses(rng, K)
    T = lambda x: torch.tensor(x, device=dev)
    args = dict(N_real=T(Nr), M_real=T(Mr), BLOCK=64, num_warps=4)
    if esp:
        V, dQ, dT = kt.overlap_score_grad_esp_se3_batch(T(A), T(B), T(cA), T(cB), T(q), T(t), alpha=0.81,
                                                        lam=0.3, **args)
    else:
        V, dQ, dT = kt.overlap_score_grad_se3_batch(T(A), T(B), T(q), T(t), alpha=0.81, **args)
    assert V.dtype == torch.float32 and V.shape == (K,) and dQ.shape == (K, 4) and dT.shape == (K, 3)
    for k in range(K):
        qk = torch.tensor(q[k:k + 1], dtype=DT)
        R = orc.rot_from_unit_quat(qk)
        w = orc.esp_pair_weights(torch.tensor(cB[k, :Mr[k]], dtype=DT), torch.tensor(cA[k, :Nr[k]], dtype=DT),
                                 0.3) if esp else None
        O, gR, gt = orc.overlap_grad_shape(R, torch.tensor(t[k:k + 1], dtype=DT),
                                           torch.tensor(A[k:k + 1, :Nr[k]], dtype=DT),
                                           torch.tensor(B[k:k + 1, :Mr[k]], dtype=DT), 0.81, w)
        gQ = orc.project_grad_R(gR, qk)
        _check(V[k:k + 1].cpu().numpy(), dQ[k:k + 1].cpu().numpy(), dT[k:k + 1].cpu().numpy(),
               O.numpy(), gQ.numpy(), gt.numpy())
    V0, dQ0, dT0 = kt.overlap_score_grad_se3_batch(T(A), T(B), T(q), T(t), N_real=T(Nr), M_real=T(Mr),
                                                   NEED_GRAD=False) if not esp else (None, None, None)
    if V0 is not None:
        np.testing.assert_array_equal(V0.cpu().numpy(), V.cpu().numpy())
        assert not dQ0.any() and not dT0.any()


def test_self_overlaps(dev):
    print("case kernel table self overlaps", flush=True)
    rng = np.random.default_rng(2)
    K, N = 4, 25
    Nr = rng.integers(5, N + 1, size=K)
    P = _padded(rng, K, N, Nr)
    C = rng.normal(0, 0.2, size=(K, N)).astype(np.float32)
    T = lambda x: torch.tensor(x, device=dev)
    S = kt._batch_self_overlap(T(P), T(Nr), alpha=0.81).cpu().numpy()
    E = kt._batch_self_overlap_esp(T(P), T(C), T(Nr), alpha=0.81, lam=0.3).cpu().numpy()
    for k in range(K):
        X = torch.tensor(P[k:k + 1, :Nr[k]], dtype=DT)
        c = torch.tensor(C[k, :Nr[k]], dtype=DT)
        np.testing.assert_allclose(S[k], float(orc.vab(X, X, 0.81)[0]), rtol=2e-5)
        np.testing.assert_allclose(E[k], float(orc.vab(X, X, 0.81, orc.esp_pair_weights(c, c, 0.3))[0]), rtol=2e-5)


def _pharm_case(rng, K, N, M):
    Nr, Mr = rng.integers(3, N + 1, size=K), rng.integers(3, M + 1, size=K)
    ta = rng.integers(0, 8, size=(K, N)); tb = rng.integers(0, 8, size=(K, M))
    tb[:, :3] = ta[:, :3]
    A = _padded(rng, K, N, np.full(K, N)); B = _padded(rng, K, M, np.full(K, M))
    va = rng.normal(size=(K, N, 3)).astype(np.float32); vb = rng.normal(size=(K, M, 3)).astype(np.float32)
    va /= np.linalg.norm(va, axis=2, keepdims=True); vb /= np.linalg.norm(vb, axis=2, keepdims=True)
    return Nr, Mr, ta, tb, A, B, va, vb


@pytest.mark.parametrize("kind", ["dq", "color", "rot"])
def test_pharm_kernels(dev, kind):
    print(f"case kernel table pharm {kind}", flush=True)
    rng = np.random.default_rng(3)
    K, N, M = 4, 10, 12
    Nr, Mr, ta, tb, A, B, va, vb = _pharm_case(rng, K, N, M)
    q, t = _poses(rng, K)
    alphas, Ks, cats = orc.pharm_tables()
    T = lambda x: torch.tensor(x, device=dev)
    tabs = (T(alphas.numpy()), T(Ks.numpy()), T(cats.numpy()))
    if kind == "dq":
        O, dQ, dT = kt.pharm_grad_dq_se3_batch(T(q), T(t), T(ta), T(tb), T(A), T(B), T(va), T(vb), *tabs,
                                               N_real=T(Nr), M_real=T(Mr))
    elif kind == "color":
        O, dQ, dT = kt.pharm_color_score_grad_se3_batch(T(A), T(B), T(q), T(t), T(ta), T(tb), *tabs,
                                                        N_real=T(Nr), M_real=T(Mr))
    else:
        Rm = orc.rot_from_unit_quat(torch.tensor(q, dtype=DT)).to(torch.float32)
        O, dR, dT = kt.pharm_score_grad_se3_batch(Rm.to(dev), T(t), T(ta), T(tb), T(A), T(B), T(va), T(vb), *tabs,
                                                  N_real=T(Nr), M_real=T(Mr))
    for k in range(K):
        qk = torch.tensor(q[k:k + 1], dtype=DT)
        R = orc.rot_from_unit_quat(qk)
        n, m = Nr[k], Mr[k]
        Oo, gR, gt = orc.overlap_grad_pharm(R, torch.tensor(t[k:k + 1], dtype=DT), torch.tensor(ta[k:k + 1, :n]),
                                            torch.tensor(tb[k:k + 1, :m]), torch.tensor(A[k:k + 1, :n], dtype=DT),
                                            torch.tensor(B[k:k + 1, :m], dtype=DT),
                                            torch.tensor(va[k:k + 1, :n], dtype=DT),
                                            torch.tensor(vb[k:k + 1, :m], dtype=DT), directionless=kind == "color")
        if kind == "rot":
            np.testing.assert_allclose(O[k].item(), float(Oo[0]), rtol=2e-5, atol=1e-6)
            scale = max(gR.abs().max().item(), gt.abs().max().item(), 1e-6)
            np.testing.assert_allclose(dR[k].cpu().numpy(), gR[0].numpy(), atol=1e-4 * scale)
            np.testing.assert_allclose(dT[k].cpu().numpy(), gt[0].numpy(), atol=1e-4 * scale)
        else:
            _check(O[k:k + 1].cpu().numpy(), dQ[k:k + 1].cpu().numpy(), dT[k:k + 1].cpu().numpy(),
                   Oo.numpy(), orc.project_grad_R(gR, qk).numpy(), gt.numpy())


def test_esp_comparison(dev):
    print("case kernel table esp comparison", flush=True)
    rng = np.random.default_rng(4)
    ms = [synth.shaep_molecule(rng, 10, 70) for _ in range(3)]
    K, N, M = 3, 70, max(len(m["centers_w_H"]) for m in ms)
    P = np.stack([m["surf_pos"] for m in ms]); PE = np.stack([m["surf_esp"] for m in ms])
    other = [ms[(k + 1) % 3] for k in range(K)]
    Mr = np.array([len(o["centers_w_H"]) for o in other])
    A = np.full((K, M, 3), 50.0, np.float32); Q = np.zeros((K, M), np.float32); Rd = np.ones((K, M), np.float32)
    for k, o in enumerate(other):
        A[k, :Mr[k]] = o["centers_w_H"] + 0.4
        Q[k, :Mr[k]] = o["partial_charges"]
        Rd[k, :Mr[k]] = o["radii"]
        A[k, Mr[k]:] = P[k, :M - Mr[k]]            # padding that would mask every point it touched
    T = lambda x: torch.tensor(x, device=dev)
    S = kt.esp_comparison_batch(T(P), T(A), T(Q), T(PE), T(Rd), M_real=T(Mr), probe_radius=1.0,
                                lam=0.001).cpu().numpy()
    for k in range(K):
        ref = orc.esp_comparison(torch.tensor(P[k:k + 1], dtype=DT), torch.tensor(A[k:k + 1, :Mr[k]], dtype=DT),
                                 torch.tensor(Q[k, :Mr[k]], dtype=DT), torch.tensor(PE[k], dtype=DT),
                                 torch.tensor(Rd[k, :Mr[k]], dtype=DT), 1.0, LAM_SCALING * 0.001)
        np.testing.assert_allclose(S[k], float(ref[0]), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("proj", [False, True])
def test_fused_adam(dev, proj):
    print(f"case kernel table adam proj={proj}", flush=True)
    rng = np.random.default_rng(5)
    n = 300
    q0, t0 = _poses(rng, n)
    state = {k: torch.zeros(s, device=dev) for k, s in (("mq", (n, 4)), ("vq", (n, 4)), ("mt", (n, 3)),
                                                        ("vt", (n, 3)))}
    ref = {k: v.clone().cpu() for k, v in state.items()}
    q, t = torch.tensor(q0, device=dev), torch.tensor(t0, device=dev)
    rq, rt = q.cpu().clone(), t.cpu().clone()
    fn = kt.fused_adam_qt_with_tangent_proj if proj else kt.fused_adam_qt
    for step in range(3):
        dQ = torch.tensor(rng.normal(size=(n, 4)).astype(np.float32))
        dT = torch.tensor(rng.normal(size=(n, 3)).astype(np.float32))
        fn(q, t, dQ.to(dev), dT.to(dev), state["mq"], state["vq"], state["mt"], state["vt"], 0.1)
        # the reference's update (cpu.py:112-138), float32 torch on the CPU
        g = dQ - rq * (dQ * rq).sum(dim=1, keepdim=True) if proj else dQ
        ref["mq"].mul_(0.9).add_(g, alpha=0.1)
        ref["vq"].mul_(0.999).addcmul_(g, g, value=1e-3)
        rq.add_(-0.1 * ref["mq"] / (ref["vq"] + 1e-8).sqrt())
        ref["mt"].mul_(0.9).add_(dT, alpha=0.1)
        ref["vt"].mul_(0.999).addcmul_(dT, dT, value=1e-3)
        rt.add_(-0.1 * ref["mt"] / (ref["vt"] + 1e-8).sqrt())
        rq.div_(rq.norm(dim=1, keepdim=True))
    np.testing.assert_allclose(q.cpu().numpy(), rq.numpy(), atol=1e-6)
    np.testing.assert_allclose(t.cpu().numpy(), rt.numpy(), atol=1e-6)
    for k in state:
        np.testing.assert_allclose(state[k].cpu().numpy(), ref[k].numpy(), rtol=1e-5, atol=1e-7)
