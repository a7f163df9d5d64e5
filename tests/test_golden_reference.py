"""The FP64 oracle against vectors of the reference's own scoring and gradient functions.

tests/golden/reference_vectors.npz holds seeded inputs and the outputs of
shepherd_score.score.analytical_gradients._torch and shepherd_score.score.{gaussian_overlap,
electrostatic_scoring, pharmacophore_scoring} run in float64 (written by
tests/golden/make_reference_vectors.py, in a container with the reference source).  Every oracle
function below must reproduce them to 1e-12: the pharmacophore category rules, extended points,
Tversky denominators, ESP pair weights, the ShaEP combination and the SE(3) loss scaling are
conventions that no closed form fixes.
"""
import os

import numpy as np
import pytest
import torch

from oracle import fss_oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
Z = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))
TOL = 1e-12
SIMS = ("tanimoto", "tversky", "tversky_ref", "tversky_fit")
GOP_VARIANTS = ((False, False, False), (True, False, False), (True, True, False), (False, False, True))
PH_VARIANTS = ((False, False), (True, False), (True, True))

T = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64)
B = Z["se3"].shape[0]
SE3 = T(Z["se3"])
R = T(Z["out_R"])
t = T(Z["t"])
e = lambda x: x.unsqueeze(0).expand(B, *x.shape)
posed = lambda x: torch.bmm(R, e(x).transpose(1, 2)).transpose(1, 2) + t[:, None, :]
rotated = lambda v: torch.bmm(R, e(v).transpose(1, 2)).transpose(1, 2)


def close(got, key, tol=TOL):
    want = Z["out_" + key]
    got = np.asarray(got.detach() if torch.is_tensor(got) else got, np.float64).reshape(want.shape)
    scale = max(1.0, float(np.abs(want).max()))
    err = float(np.abs(got - want).max()) / scale
    assert err <= tol, (key, err)


def pharm_pair(ext=False, only=False, similarity="tanimoto", directionless=False):
    return orc.Pair("pharm", ref_types=Z["ref_typ"], fit_types=Z["fit_typ"], ref_anc=Z["ref_anc"],
                    fit_anc=Z["fit_anc"], ref_vec=Z["ref_vec"], fit_vec=Z["fit_vec"], similarity=similarity,
                    directionless=directionless, extended_points=ext, only_extended=only)


def test_rotation_of_unit_quaternions():
    q = SE3[:, :4] / torch.norm(SE3[:, :4], dim=1, keepdim=True)
    close(orc.rot_from_unit_quat(q), "R")


def test_shape_overlap_gradients_and_se3_loss():
    ref, fit, a = T(Z["ref"]), T(Z["fit"]), float(Z["alpha"])
    O, gR, gt = orc.overlap_grad_shape(R, t, e(ref), e(fit), a)
    close(O, "shape_O"); close(gR, "shape_gR"); close(gt, "shape_gt")
    p = orc.Pair("shape", ref=Z["ref"], fit=Z["fit"], alpha=a)
    so = p.self_overlaps()
    close(so["VAA_s"], "shape_VAA"); close(so["VBB_s"], "shape_VBB")
    loss, g, _ = p.loss_grad_se3(SE3, so)
    close(loss, "shape_loss"); close(g, "shape_gse3")
    close(orc.tanimoto_shape(e(ref), posed(fit), a), "shape_tanimoto")


def test_esp_weights_overlap_gradients_and_se3_loss():
    ref, fit, a, lam = T(Z["ref"]), T(Z["fit"]), float(Z["alpha"]), float(Z["out_esp_lam"])
    qr, qf = T(Z["ref_esp"]), T(Z["fit_esp"])
    w = orc.esp_pair_weights(qf, qr, lam)
    close(w, "esp_w")
    O, gR, gt = orc.overlap_grad_shape(R, t, e(ref), e(fit), a, w)
    close(O, "esp_O"); close(gR, "esp_gR"); close(gt, "esp_gt")
    p = orc.Pair("esp", ref=Z["ref"], fit=Z["fit"], ref_esp=Z["ref_esp"], fit_esp=Z["fit_esp"], alpha=a, lam=lam)
    so = p.self_overlaps()
    close(so["VAA_s"], "esp_VAA"); close(so["VBB_s"], "esp_VBB")
    loss, g, _ = p.loss_grad_se3(SE3, so)
    close(loss, "esp_loss"); close(g, "esp_gse3")
    close(orc.tanimoto_esp(e(ref), posed(fit), e(qr), e(qf), a, lam), "esp_tanimoto")


@pytest.mark.parametrize("k", range(len(PH_VARIANTS)))
def test_pharm_overlap_gradients_and_se3_loss(k):
    ext, only = PH_VARIANTS[k]
    I = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.long)
    O, gR, gt = orc.overlap_grad_pharm(R, t, e(I(Z["ref_typ"])), e(I(Z["fit_typ"])), e(T(Z["ref_anc"])),
                                       e(T(Z["fit_anc"])), e(T(Z["ref_vec"])), e(T(Z["fit_vec"])),
                                       extended_points=ext, only_extended=only)
    close(O, f"ph{k}_O"); close(gR, f"ph{k}_gR"); close(gt, f"ph{k}_gt")
    for sim in ("tanimoto", "tversky"):
        p = pharm_pair(ext, only, sim)
        so = p.self_overlaps()
        close(so["VAA_p"], f"ph{k}_VAA"); close(so["VBB_p"], f"ph{k}_VBB")
        loss, g, _ = p.loss_grad_se3(SE3, so)
        close(loss, f"ph{k}_{sim}_loss"); close(g, f"ph{k}_{sim}_gse3")


@pytest.mark.parametrize("k", range(len(GOP_VARIANTS)))
@pytest.mark.parametrize("sim", SIMS)
def test_pharm_similarity_scoring(k, sim):
    ext, only, dirless = GOP_VARIANTS[k]
    I = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.long)
    got = orc.pharm_similarity(I(Z["ref_typ"]), I(Z["fit_typ"]), T(Z["ref_anc"]), posed(T(Z["fit_anc"])),
                               T(Z["ref_vec"]), rotated(T(Z["fit_vec"])), sim, directionless=dirless,
                               extended_points=ext, only_extended=only)
    # The reference's batched extended-point overlaps multiply the Gaussian prefactor by a mask
    # cast to float32 (gaussian_overlap.py:197-198, 217-221), which rounds the prefactor to float32
    # (relative error up to 6e-8): that path agrees to 1e-7, the single-instance one to 1e-12.
    tol = 1e-7 if ext else TOL
    close(got, f"gop{k}_{sim}", tol)
    # the optimiser's per-pose score takes the same path through Pair.score_poses
    close(pharm_pair(ext, only, sim, dirless).score_poses(SE3), f"gop{k}_{sim}", tol)
    # the reference's single-instance path, on unit vectors
    unit = lambda v: v / torch.norm(v, dim=-1, keepdim=True)
    got = orc.pharm_similarity(I(Z["ref_typ"]), I(Z["fit_typ"]), T(Z["ref_anc"]), posed(T(Z["fit_anc"])),
                               unit(T(Z["ref_vec"])), unit(rotated(T(Z["fit_vec"]))), sim,
                               directionless=dirless, extended_points=ext, only_extended=only)
    close(got, f"gopu{k}_{sim}")


def test_directionless_analytic_overlap_scores_like_the_reference():
    """The optimiser's directionless (color) overlap: its Tanimoto equals the reference's
    directionless get_overlap_pharm."""
    p = pharm_pair(directionless=True)
    so = p.self_overlaps()
    O = p.channel_overlaps(R, t)["pharm"][0]
    close(O / (so["VAA_p"] + so["VBB_p"] - O), "gop3_tanimoto")


@pytest.mark.parametrize("name", ["heavy", "surf"])
def test_shaep_combination_score(name):
    alpha = 0.81 if name == "heavy" else float(Z["alpha_surf"])
    c = lambda k: Z["combo_" + k]
    p = orc.ComboPair(ref_centers=c("heavy1") if name == "heavy" else c("surf1"),
                      fit_centers=c("heavy2") if name == "heavy" else c("surf2"),
                      ref_surf=c("surf1"), fit_surf=c("surf2"), ref_surf_esp=c("pesp1"), fit_surf_esp=c("pesp2"),
                      ref_atoms=c("atoms1"), fit_atoms=c("atoms2"), ref_charges=c("q1"), fit_charges=c("q2"),
                      ref_radii=c("rad1"), fit_radii=c("rad2"), alpha=alpha, lam=0.001, probe_radius=1.0,
                      esp_weight=0.5)
    close(p.score_poses(SE3), f"combo_{name}")


def test_vectors_exercise_every_branch():
    """The inputs reach the branches the pins are for: every pharmacophore type on both sides,
    ShaEP surface points both masked and not, ESP comparisons neither all 0 nor all 1."""
    assert set(Z["ref_typ"]) == set(range(9)) and set(Z["fit_typ"]) == set(range(9))
    d = np.linalg.norm(Z["combo_surf1"][:, None] - Z["combo_atoms2"][None], axis=-1)
    inside = (d < Z["combo_rad2"][None] + 1.0).any(1)
    assert 0 < inside.sum() < len(inside)
    # extended points, only-extended and directionless each change the score
    for k in (1, 2, 3):
        assert np.abs(Z[f"out_gop{k}_tanimoto"] - Z["out_gop0_tanimoto"]).max() > 1e-3
