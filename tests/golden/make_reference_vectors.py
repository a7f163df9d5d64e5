"""Write FP64 input/output vectors of the reference's own scoring and gradient functions.

Run in a container that has the reference source (never on a GPU machine, never from a test):
    python tests/golden/make_reference_vectors.py /root/reference

It imports only ``shepherd_score.score.analytical_gradients._torch`` and the three
``shepherd_score.score`` modules (their import chains need torch, numpy and scipy only), runs
them in float64 on small seeded inputs, and stores the inputs and outputs in
``tests/golden/reference_vectors.npz`` (plain numpy arrays, no pickles).  The oracle
(oracle/fss_oracle.py) is checked against this file by tests/test_golden_reference.py.

These are the functions whose conventions no closed form pins: which pharmacophore types take a
cosine weight and which an absolute one, the extended-point rules, the Tversky denominators, the
ESP pair weights and the ShaEP combination of masked potential comparisons with the shape
Tanimoto, and the SE(3) loss with its 1/B scaling.  Vectors written (per group, see the keys):

  R         _rotation_matrix_from_unit_quat (ag/_torch.py:28-59) of the normalised poses
  shape_*   compute_overlap_and_grad_shape (ag/_torch.py:441-516), compute_self_overlaps_shape
            (:519-541), compute_analytical_grad_se3_shape (:696-772), get_overlap
            (gaussian_overlap.py:73-106)
  esp_*     _compute_esp_pair_weights / compute_self_overlaps_esp (:544-599),
            compute_overlap_and_grad_shape with the pair weights, compute_analytical_grad_se3_esp
            (:602-693), get_overlap_esp (electrostatic_scoring.py:69-125)
  ph{e}_*   compute_overlap_and_grad_pharm (:223-361) and compute_self_overlaps_pharm (:364-389)
            for e = 0 (cosine weights), 1 (extended points), 2 (only extended points);
            compute_analytical_grad_se3 (:939-1017) with Tanimoto and Tversky (sigma 0.95)
  gop{k}_*  get_overlap_pharm (pharmacophore_scoring.py:422-912), batched over the poses, for
            every similarity x k in {plain, extended, only extended, directionless}
  gopu{k}_* the same one pose at a time (the single-instance path) on unit vectors
  combo_*   esp_combo_score (electrostatic_scoring.py:230-320) over batched poses, with both
            atom-centre and surface-point shape terms
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_vectors.npz")

SIMS = ("tanimoto", "tversky", "tversky_ref", "tversky_fit")
GOP_VARIANTS = ((False, False, False), (True, False, False), (True, True, False), (False, False, True))


def inputs(seed=20261204):
    """Seeded inputs of every group (float64)."""
    rng = np.random.default_rng(seed)
    d = {}
    B = 4
    # poses: raw (unnormalised) quaternions + translations (the rotation matrices of the unit
    # quaternions come from the reference's own function, in main)
    q_raw = rng.normal(size=(B, 4)) * np.array([[1.5, 0.7, 0.7, 0.7]])
    d["se3"] = np.concatenate([q_raw, rng.normal(scale=0.8, size=(B, 3))], 1)
    d["t"] = d["se3"][:, 4:].copy()
    # shape / ESP clouds
    d["ref"] = rng.normal(scale=1.6, size=(11, 3))
    d["fit"] = rng.normal(scale=1.6, size=(8, 3))
    d["ref_esp"] = rng.normal(scale=6.0, size=11)
    d["fit_esp"] = rng.normal(scale=6.0, size=8)
    d["alpha"] = np.array(0.81)
    d["alpha_surf"] = np.array(1.37)
    # pharmacophores: every type (dummy included) on both sides, several of each
    d["ref_typ"] = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 0, 2, 3, 1, 4], np.int64)
    d["fit_typ"] = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 2, 0, 1, 3], np.int64)
    d["ref_anc"] = rng.normal(scale=1.5, size=(len(d["ref_typ"]), 3))
    d["fit_anc"] = rng.normal(scale=1.5, size=(len(d["fit_typ"]), 3))
    # vectors of assorted lengths (the functions normalise them)
    d["ref_vec"] = rng.normal(size=(len(d["ref_typ"]), 3)) * rng.uniform(0.5, 2.0, size=(len(d["ref_typ"]), 1))
    d["fit_vec"] = rng.normal(size=(len(d["fit_typ"]), 3)) * rng.uniform(0.5, 2.0, size=(len(d["fit_typ"]), 1))
    # ShaEP: atoms with hydrogens, heavy-atom centres, surface points with ESP, radii
    for side, na, nh, ns in (("1", 12, 7, 30), ("2", 10, 6, 26)):
        atoms = rng.normal(scale=1.3, size=(na, 3))
        d["combo_atoms" + side] = atoms
        d["combo_heavy" + side] = atoms[:nh].copy()
        d["combo_q" + side] = rng.normal(scale=0.3, size=na)
        d["combo_rad" + side] = rng.uniform(1.1, 1.9, size=na)
        u = rng.normal(size=(ns, 3))
        d["combo_surf" + side] = u / np.linalg.norm(u, axis=1, keepdims=True) * rng.uniform(2.5, 5.0, size=(ns, 1))
    return d


def combo_point_esp(d):
    """Surface ESP values near each side's potential from the other side's atoms (so that the
    masked comparisons are neither all 0 nor all 1): the potential from the other molecule at
    the identity pose plus noise."""
    rng = np.random.default_rng(7)
    C = 1e4 / (4 * 55.263 * np.pi)
    out = {}
    for a, b in (("1", "2"), ("2", "1")):
        p, x, q = d["combo_surf" + a], d["combo_atoms" + b], d["combo_q" + b]
        phi = C * (q[None, :] / np.linalg.norm(p[:, None] - x[None], axis=-1)).sum(1)
        out["combo_pesp" + a] = phi + rng.normal(scale=0.25, size=len(p))
    return out


def main(ref_root: str) -> None:
    import torch
    torch.set_default_dtype(torch.float64)
    sys.path.insert(0, ref_root)
    from shepherd_score.score import electrostatic_scoring as es
    from shepherd_score.score import gaussian_overlap as go
    from shepherd_score.score import pharmacophore_scoring as ps
    from shepherd_score.score.analytical_gradients import _torch as ag
    from shepherd_score.score.constants import LAM_SCALING

    d = inputs()
    d.update(combo_point_esp(d))
    T = lambda a: torch.as_tensor(a, dtype=torch.float64)
    I = lambda a: torch.as_tensor(a, dtype=torch.long)
    out = {}
    B = d["se3"].shape[0]
    se3, t = T(d["se3"]), T(d["t"])
    R = ag._rotation_matrix_from_unit_quat(se3[:, :4] / torch.norm(se3[:, :4], dim=1, keepdim=True))
    out["R"] = R
    ref, fit = T(d["ref"]), T(d["fit"])
    alpha = float(d["alpha"])
    e = lambda x: x.unsqueeze(0).expand(B, *x.shape).contiguous()
    posed = lambda x: (torch.bmm(R, e(x).transpose(1, 2)).transpose(1, 2) + t[:, None, :]).contiguous()
    rotated = lambda v: torch.bmm(R, e(v).transpose(1, 2)).transpose(1, 2).contiguous()

    # ---- shape
    O, gR, gt = ag.compute_overlap_and_grad_shape(R, t, e(ref), e(fit), alpha)
    VAA, VBB = ag.compute_self_overlaps_shape(ref, fit, alpha)
    loss, g = ag.compute_analytical_grad_se3_shape(se3, e(ref), e(fit), alpha, VAA, VBB)
    out.update(shape_O=O, shape_gR=gR, shape_gt=gt, shape_VAA=VAA, shape_VBB=VBB, shape_loss=loss,
               shape_gse3=g, shape_tanimoto=go.get_overlap(e(ref), posed(fit), alpha))
    # ---- ESP (surface convention lam = 0.3 * LAM_SCALING)
    lam = 0.3 * LAM_SCALING
    qr, qf = T(d["ref_esp"]), T(d["fit_esp"])
    w = ag._compute_esp_pair_weights(qf, qr, lam)
    O, gR, gt = ag.compute_overlap_and_grad_shape(R, t, e(ref), e(fit), alpha, pair_weights=w)
    VAA, VBB = ag.compute_self_overlaps_esp(ref, fit, qr, qf, alpha, lam)
    loss, g = ag.compute_analytical_grad_se3_esp(se3, e(ref), e(fit), qr, qf, alpha, lam, VAA, VBB)
    out.update(esp_lam=torch.tensor(lam), esp_w=w, esp_O=O, esp_gR=gR, esp_gt=gt, esp_VAA=VAA, esp_VBB=VBB,
               esp_loss=loss, esp_gse3=g,
               esp_tanimoto=es.get_overlap_esp(e(ref), posed(fit), e(qr), e(qf), alpha, lam))
    # ---- pharmacophores, analytic path
    rt, ft = I(d["ref_typ"]), I(d["fit_typ"])
    ra, fa, rv, fv = T(d["ref_anc"]), T(d["fit_anc"]), T(d["ref_vec"]), T(d["fit_vec"])
    for k, (ext, only) in enumerate(((False, False), (True, False), (True, True))):
        O, gR, gt = ag.compute_overlap_and_grad_pharm(R, t, e(rt), e(ft), e(ra), e(fa), e(rv), e(fv),
                                                      extended_points=ext, only_extended=only)
        VAA, VBB = ag.compute_self_overlaps_pharm(rt, ft, ra, fa, rv, fv, extended_points=ext, only_extended=only)
        out.update({f"ph{k}_O": O, f"ph{k}_gR": gR, f"ph{k}_gt": gt, f"ph{k}_VAA": VAA, f"ph{k}_VBB": VBB})
        for sim, sigma in (("tanimoto", 0.5), ("tversky", 0.95)):
            loss, g = ag.compute_analytical_grad_se3(se3, e(rt), e(ft), e(ra), e(fa), e(rv), e(fv), VAA, VBB,
                                                     similarity=sim, sigma=sigma, extended_points=ext,
                                                     only_extended=only)
            out.update({f"ph{k}_{sim}_loss": loss, f"ph{k}_{sim}_gse3": g})
    # ---- pharmacophores, scoring path: the batched call (vectors of any length, normalised by
    # the function), and one pose at a time on unit vectors (the single-instance path adds the
    # vectors to the anchors as given, so it agrees with the batched one for unit vectors only)
    pa, pv = posed(fa), rotated(fv)
    unit = lambda v: v / torch.norm(v, dim=-1, keepdim=True)
    for k, (ext, only, dirless) in enumerate(GOP_VARIANTS):
        for sim in SIMS:
            out[f"gop{k}_{sim}"] = ps.get_overlap_pharm(e(rt), e(ft), e(ra), pa, e(rv), pv, similarity=sim,
                                                        extended_points=ext, only_extended=only,
                                                        directionless=dirless).reshape(B)
            s = [ps.get_overlap_pharm(rt, ft, ra, pa[b], unit(rv), unit(pv[b]), similarity=sim,
                                      extended_points=ext, only_extended=only, directionless=dirless)
                 for b in range(B)]
            out[f"gopu{k}_{sim}"] = torch.stack([torch.as_tensor(x).reshape(()) for x in s])
    # ---- ShaEP combination: heavy-atom shape (alpha 0.81) and surface shape (alpha_surf)
    pc = {k: T(d[k]) for k in d if k.startswith("combo_")}
    x1, x2 = pc["combo_atoms1"], posed(pc["combo_atoms2"])
    s1, s2 = pc["combo_surf1"], posed(pc["combo_surf2"])
    for name, c1, c2, a in (("heavy", pc["combo_heavy1"], posed(pc["combo_heavy2"]), 0.81),
                            ("surf", pc["combo_surf1"], s2, float(d["alpha_surf"]))):
        out[f"combo_{name}"] = es.esp_combo_score(
            e(x1), x2, e(c1), c2, e(s1), s2, e(pc["combo_q1"]), e(pc["combo_q2"]),
            e(pc["combo_pesp1"]), e(pc["combo_pesp2"]), e(pc["combo_rad1"]), e(pc["combo_rad2"]),
            alpha=a, lam=0.001, probe_radius=1.0, esp_weight=0.5).reshape(B)
    arrays = {k: np.asarray(v) for k, v in d.items()}
    arrays.update({"out_" + k: np.asarray(v.detach().cpu().numpy(), np.float64) for k, v in out.items()})
    np.savez(OUT, **arrays)
    print(f"wrote {OUT}: {len(arrays)} arrays")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
