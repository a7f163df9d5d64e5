"""FP64 CPU oracle of the ShEPhERD scoring + alignment hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import this module, and only to check results against -- never as the
thing measured or shipped.  The product (``fast_shepherd_score_amd``) never imports it.

A restatement, in torch float64 on the CPU, of:
  * score/analytical_gradients/_torch.py:28-1017  (R(q), dR/dq, projection, overlap + analytic
    gradients for shape / ESP / pharmacophore, self-overlaps, Tanimoto / Tversky chain rules,
    the batched SE(3) loss + gradient with its 1/B scaling);
  * alignment/_torch_analytical.py:23-478 (the optimisers: torch.optim.Adam on raw (q, t), the
    |dL| <= 1e-5 for > 10 steps early stop on the restart-mean loss, final re-scoring, argmax);
  * alignment/_torch.py:1296-1548 (the shape + pharmacophore combination, objective 1 - mean of
    (1-w) shape Tanimoto + w pharm similarity, restated with the analytic gradients above);
  * the ShaEP combination score/electrostatic_scoring.py:128-320 with its autograd optimiser
    alignment/_torch.py:804-1002 (ComboPair);
  * the final re-scoring functions score/gaussian_overlap.py:16-106,
    score/electrostatic_scoring.py:17-125, score/pharmacophore_scoring.py:123-145, 422-912,
    score/gaussian_overlap.py:283-363 and alignment/utils/se3.py:16-189, restated separately
    from the analytic overlap so the two restatements check each other.

Differences from the shipped reference, by design: everything is float64 (the reference
optimisers allocate float32 seeds and matrices, alignment/_torch.py:292,297, se3.py:127,143, so
they cannot run in FP64 as shipped), and restart seeds are passed in explicitly.

Pinning (tests/test_oracle.py): closed forms, brute-force double loops, finite differences,
torch.autograd of the scoring restatement, the reference's known-answer rotation
(tests/test_alignment_utils.py:50-56), rigid-motion invariances.  tests/test_golden_reference.py
checks the overlap / gradient / SE(3) loss functions, the pharmacophore, ESP and ShaEP scores
against float64 outputs of the reference's own score modules (tests/golden/reference_vectors.npz)
to 1e-12.  The absolute optimiser trajectory (torch.optim.Adam) is pinned only through the
library routine itself: parity unpinned beyond that, as the reference holds no golden optimiser
outputs.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch
import torch.nn.functional as F

DT = torch.float64
LAM_SCALING = (1e4 / (4 * 55.263 * np.pi)) ** 2          # score/constants.py:16-17
P_TYPES = ("acceptor", "donor", "aromatic", "hydrophobe", "halogen", "cation", "anion",
           "znbinder", "dummy")                             # score/constants.py:20
P_ALPHAS = {"acceptor": 1.0, "donor": 1.0, "aromatic": 0.7, "hydrophobe": 0.7, "cation": 1.0,
            "anion": 1.0, "znbinder": 1.0, "halogen": 1.0, "dummy": 1.0}   # constants.py:24-34
SIGMA = {"tversky": 0.95, "tversky_ref": 1.0, "tversky_fit": 0.05}


def K_of(alpha: float) -> float:
    return (math.pi / (2.0 * alpha)) ** 1.5


# ------------------------------------------------------------------------------------------
# Rotations (ag/_torch.py:28-173)
# ------------------------------------------------------------------------------------------
def rot_from_unit_quat(q: torch.Tensor) -> torch.Tensor:
    """(B,4) unit (w,x,y,z) -> (B,3,3).  ag/_torch.py:28-59."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
                       dim=-1).reshape(-1, 3, 3)


def rot_jacobians(q: torch.Tensor):
    """dR/dq_k for k = w, x, y, z at unit q, each (B,3,3).  ag/_torch.py:62-139."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    o = torch.zeros_like(w)
    Jw = torch.stack([o, -2 * z, 2 * y, 2 * z, o, -2 * x, -2 * y, 2 * x, o], -1).reshape(-1, 3, 3)
    Jx = torch.stack([o, 2 * y, 2 * z, 2 * y, -4 * x, -2 * w, 2 * z, 2 * w, -4 * x], -1).reshape(-1, 3, 3)
    Jy = torch.stack([-4 * y, 2 * x, 2 * w, 2 * x, o, 2 * z, -2 * w, 2 * z, -4 * y], -1).reshape(-1, 3, 3)
    Jz = torch.stack([-4 * z, -2 * w, 2 * x, 2 * w, -4 * z, 2 * y, 2 * x, 2 * y, o], -1).reshape(-1, 3, 3)
    return Jw, Jx, Jy, Jz


def project_grad_R(G: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
    """dL/dq_k = sum_ij G_ij (dR/dq_k)_ij.  ag/_torch.py:142-173."""
    return torch.stack([(G * J).sum(dim=(-2, -1)) for J in rot_jacobians(q)], dim=-1)


def se3_matrix(se3: torch.Tensor) -> torch.Tensor:
    """(B,7) -> (B,4,4) with q normalised (F.normalize) and R from 2/|q|^2.  se3.py:32-149."""
    q = F.normalize(se3[:, :4], p=2, dim=1)
    r, i, j, k = q.unbind(1)
    two_s = 2.0 / (q ** 2).sum(1)
    R = torch.stack([1 - two_s * (j * j + k * k), two_s * (i * j - k * r), two_s * (i * k + j * r),
                     two_s * (i * j + k * r), 1 - two_s * (i * i + k * k), two_s * (j * k - i * r),
                     two_s * (i * k - j * r), two_s * (j * k + i * r), 1 - two_s * (i * i + j * j)],
                    1).reshape(-1, 3, 3)
    T = torch.eye(4, dtype=se3.dtype).repeat(se3.shape[0], 1, 1)
    T[:, :3, :3] = R
    T[:, :3, 3] = se3[:, 4:]
    return T


def apply_se3(points: torch.Tensor, T: torch.Tensor) -> torch.Tensor:
    """points (B,N,3), T (B,4,4) -> points @ R^T + t.  se3.py:152-189."""
    return torch.baddbmm(T[:, :3, 3].unsqueeze(1), points, T[:, :3, :3].transpose(1, 2))


def apply_so3(vectors: torch.Tensor, T: torch.Tensor) -> torch.Tensor:
    """Rotation only.  se3.py:192-230."""
    return torch.bmm(T[:, :3, :3], vectors.permute(0, 2, 1)).permute(0, 2, 1)


# ------------------------------------------------------------------------------------------
# Overlaps and analytic gradients (ag/_torch.py:176-599)
# ------------------------------------------------------------------------------------------
def pharm_tables(directionless: bool = False):
    """(alphas, Ks, categories) in P_TYPES order.  ag/_torch.py:176-221."""
    nondir = {"hydrophobe", "znbinder", "anion", "cation"}
    dirn = {"acceptor", "donor", "halogen"}
    alphas, Ks, cats = [], [], []
    for p in P_TYPES:
        a = P_ALPHAS[p]
        alphas.append(a)
        Ks.append(K_of(a))
        if p == "dummy":
            cats.append(3)
        elif directionless or p in nondir:
            cats.append(0)
        elif p in dirn:
            cats.append(1)
        elif p == "aromatic":
            cats.append(2)
        else:
            cats.append(3)
    return (torch.tensor(alphas, dtype=DT), torch.tensor(Ks, dtype=DT),
            torch.tensor(cats, dtype=torch.long))


def overlap_grad_shape(R, t, ref, fit, alpha, pair_weights=None):
    """O_AB, dO/dR, dO/dt for batched poses.  ag/_torch.py:441-516.

    R (B,3,3), t (B,3), ref (B,N,3), fit (B,M,3), pair_weights (M,N) or (B,M,N)."""
    K = K_of(alpha)
    fit_t = torch.bmm(R, fit.permute(0, 2, 1)).permute(0, 2, 1) + t.unsqueeze(1)
    d2 = ((fit_t.unsqueeze(2) - ref.unsqueeze(1)) ** 2).sum(-1)          # (B, M, N)
    E = torch.exp(-alpha / 2.0 * d2)
    if pair_weights is not None:
        E = E * pair_weights
    O = K * E.sum(dim=(1, 2))
    aKE = -alpha * K * E
    s_ref = aKE.sum(dim=2)
    s_fit = aKE.sum(dim=1)
    grad_t = (s_ref.unsqueeze(-1) * fit_t).sum(1) - (s_fit.unsqueeze(-1) * ref).sum(1)
    delta = s_ref.unsqueeze(-1) * fit_t - torch.bmm(aKE, ref)
    grad_R = torch.bmm(delta.transpose(1, 2), fit)
    return O, grad_R, grad_t


def esp_pair_weights(c_fit, c_ref, lam):
    """w_ab = exp(-(c_fit_b - c_ref_a)^2 / lam), (M,N).  ag/_torch.py:544-566."""
    return torch.exp(-((c_fit.unsqueeze(-1) - c_ref.unsqueeze(-2)) ** 2) / lam)


def overlap_grad_pharm(R, t, rtyp, ftyp, ranc, fanc, rvec, fvec, directionless=False,
                       extended_points=False, only_extended=False):
    """Typed directional pharmacophore overlap and gradients.  ag/_torch.py:223-361.
    All inputs batched (B, ...).  ``extended_points``: directional (category 1) pairs lose
    their cosine weight (anchor weight 1, or 0 with ``only_extended``) and add an isotropic
    overlap of the extended points anchor + unit vector (ag/_torch.py:288-297, 319-357)."""
    rtyp = rtyp.long()
    ftyp = ftyp.long()
    fanc_t = torch.bmm(R, fanc.permute(0, 2, 1)).permute(0, 2, 1) + t.unsqueeze(1)
    rvn = F.normalize(rvec, p=2, dim=-1)
    fvn0 = F.normalize(fvec, p=2, dim=-1)
    fvn = torch.bmm(R, fvn0.permute(0, 2, 1)).permute(0, 2, 1)
    alphas, Ks, cats = pharm_tables(directionless)
    mask = (ftyp.unsqueeze(2) == rtyp.unsqueeze(1)) & (cats[ftyp] != 3).unsqueeze(2)
    a_ab = alphas[ftyp].unsqueeze(2)
    K_ab = Ks[ftyp].unsqueeze(2)
    c_ab = cats[ftyp].unsqueeze(2)
    d2 = ((fanc_t.unsqueeze(2) - ranc.unsqueeze(1)) ** 2).sum(-1)
    E = torch.exp(-a_ab / 2.0 * d2) * mask.to(DT)
    D = torch.bmm(fvn, rvn.permute(0, 2, 1))
    w_dir = (D.clamp(0.0, 1.0) + 2.0) / 3.0
    if extended_points:
        w_dir = torch.full_like(D, 0.0 if only_extended else 1.0)
    w = torch.where(c_ab == 1, w_dir,
                    torch.where(c_ab == 2, (D.abs() + 2.0) / 3.0, torch.ones_like(D)))
    wE = w * E
    O = (K_ab * wE).sum(dim=(1, 2))
    aKwE = -a_ab * K_ab * wE
    s_ref = aKwE.sum(2)
    s_fit = aKwE.sum(1)
    grad_t = (s_ref.unsqueeze(-1) * fanc_t).sum(1) - (s_fit.unsqueeze(-1) * ranc).sum(1)
    delta = s_ref.unsqueeze(-1) * fanc_t - torch.bmm(aKwE, ranc)
    grad_R = torch.bmm(delta.transpose(1, 2), fanc)
    cdir = ((D > 0.0) & (D < 1.0)).to(DT)
    if extended_points:
        cdir = torch.zeros_like(D)             # constant anchor weight: no weight gradient
    carom = torch.sign(D)
    c = torch.where(c_ab == 1, cdir, torch.where(c_ab == 2, carom, torch.zeros_like(D)))
    coeff = (1.0 / 3.0) * K_ab * E * c
    grad_R = grad_R + torch.bmm(torch.bmm(rvn.transpose(1, 2), coeff.transpose(1, 2)), fvn0)
    if extended_points:
        # the fit extended point moves as R (p + u0) + t
        fext = fanc_t + fvn
        rext = ranc + rvn
        Ee = torch.exp(-a_ab / 2.0 * ((fext.unsqueeze(2) - rext.unsqueeze(1)) ** 2).sum(-1))
        Ee = Ee * (mask & (c_ab == 1)).to(DT)
        O = O + (K_ab * Ee).sum(dim=(1, 2))
        aKE = -a_ab * K_ab * Ee
        se_ref = aKE.sum(2)
        se_fit = aKE.sum(1)
        grad_t = grad_t + (se_ref.unsqueeze(-1) * fext).sum(1) - (se_fit.unsqueeze(-1) * rext).sum(1)
        de = se_ref.unsqueeze(-1) * fext - torch.bmm(aKE, rext)
        grad_R = grad_R + torch.bmm(de.transpose(1, 2), fanc + fvn0)
    return O, grad_R, grad_t


def tanimoto_chain(O, U, gR, gt):
    """loss = 1 - O/(U-O); scaled grads.  ag/_torch.py:392-413."""
    den = U - O
    S = O / den
    scale = -U / (den * den)
    return 1.0 - S, scale[:, None, None] * gR, scale[:, None] * gt


def tversky_chain(O, Dn, gR, gt):
    """loss = 1 - min(O/D, 1).  ag/_torch.py:416-438."""
    S = torch.clamp_max(O / Dn, 1.0)
    scale = -(O < Dn).to(DT) / Dn
    return 1.0 - S, scale[:, None, None] * gR, scale[:, None] * gt


# ------------------------------------------------------------------------------------------
# One pair's objective
# ------------------------------------------------------------------------------------------
@dataclass
class Pair:
    """One (ref, fit) pair and its objective.

    mode: 'shape' | 'esp' | 'pharm' | 'shape_pharm'.  Shape arrays (N,3)/(M,3), ESP values
    (N,)/(M,), pharm types (n,)/(m,), anchors and vectors (n,3)/(m,3)."""
    mode: str
    ref: Optional[np.ndarray] = None
    fit: Optional[np.ndarray] = None
    ref_esp: Optional[np.ndarray] = None
    fit_esp: Optional[np.ndarray] = None
    ref_types: Optional[np.ndarray] = None
    fit_types: Optional[np.ndarray] = None
    ref_anc: Optional[np.ndarray] = None
    fit_anc: Optional[np.ndarray] = None
    ref_vec: Optional[np.ndarray] = None
    fit_vec: Optional[np.ndarray] = None
    alpha: float = 0.81
    lam: float = 0.3 * LAM_SCALING
    color_weight: float = 0.5
    similarity: str = "tanimoto"
    directionless: bool = False
    extended_points: bool = False
    only_extended: bool = False

    def __post_init__(self):
        f = lambda a: None if a is None else torch.as_tensor(np.asarray(a, np.float64), dtype=DT)
        self.tref, self.tfit = f(self.ref), f(self.fit)
        self.tref_esp, self.tfit_esp = f(self.ref_esp), f(self.fit_esp)
        self.tref_anc, self.tfit_anc = f(self.ref_anc), f(self.fit_anc)
        self.tref_vec, self.tfit_vec = f(self.ref_vec), f(self.fit_vec)
        g = lambda a: None if a is None else torch.as_tensor(np.asarray(a, np.int64))
        self.tref_typ, self.tfit_typ = g(self.ref_types), g(self.fit_types)
        self.has_shape = self.mode in ("shape", "esp", "shape_pharm")
        self.has_pharm = self.mode in ("pharm", "shape_pharm")
        if self.directionless:     # pharmacophore_scoring.py:517-523
            self.extended_points = self.only_extended = False
        self.ph_kw = dict(directionless=self.directionless, extended_points=self.extended_points,
                          only_extended=self.only_extended)

    # -- self-overlaps (ag/_torch.py:364-389, 519-541, 569-599)
    def self_overlaps(self):
        I = torch.eye(3, dtype=DT)[None]
        z = torch.zeros((1, 3), dtype=DT)
        out = {}
        if self.has_shape:
            wa = wb = None
            if self.mode == "esp":
                wa = esp_pair_weights(self.tref_esp, self.tref_esp, self.lam)
                wb = esp_pair_weights(self.tfit_esp, self.tfit_esp, self.lam)
            out["VAA_s"] = overlap_grad_shape(I, z, self.tref[None], self.tref[None], self.alpha, wa)[0][0]
            out["VBB_s"] = overlap_grad_shape(I, z, self.tfit[None], self.tfit[None], self.alpha, wb)[0][0]
        if self.has_pharm:
            out["VAA_p"] = overlap_grad_pharm(I, z, self.tref_typ[None], self.tref_typ[None], self.tref_anc[None],
                                              self.tref_anc[None], self.tref_vec[None], self.tref_vec[None],
                                              **self.ph_kw)[0][0]
            out["VBB_p"] = overlap_grad_pharm(I, z, self.tfit_typ[None], self.tfit_typ[None], self.tfit_anc[None],
                                              self.tfit_anc[None], self.tfit_vec[None], self.tfit_vec[None],
                                              **self.ph_kw)[0][0]
        return out

    def channel_overlaps(self, R, t):
        """{'shape': (O, gR, gt), 'pharm': (O, gR, gt)} for poses R (B,3,3), t (B,3)."""
        B = R.shape[0]
        res = {}
        if self.has_shape:
            w = None
            if self.mode == "esp":
                w = esp_pair_weights(self.tfit_esp, self.tref_esp, self.lam)
            res["shape"] = overlap_grad_shape(R, t, self.tref.expand(B, -1, -1), self.tfit.expand(B, -1, -1),
                                              self.alpha, w)
        if self.has_pharm:
            e = lambda a: a.expand(B, *a.shape)
            res["pharm"] = overlap_grad_pharm(R, t, e(self.tref_typ), e(self.tfit_typ), e(self.tref_anc),
                                              e(self.tfit_anc), e(self.tref_vec), e(self.tfit_vec),
                                              **self.ph_kw)
        return res

    def loss_grad_Rt(self, R, t, so):
        """Per-restart loss (B,), dL/dR (B,3,3), dL/dt (B,3)."""
        ch = self.channel_overlaps(R, t)
        loss = torch.zeros(R.shape[0], dtype=DT)
        gR = torch.zeros_like(R)
        gt = torch.zeros_like(t)
        ws = (1.0 - self.color_weight) if self.mode == "shape_pharm" else 1.0
        wp = self.color_weight if self.mode == "shape_pharm" else 1.0
        if self.has_shape:
            O, a, b = ch["shape"]
            l, a, b = tanimoto_chain(O, so["VAA_s"] + so["VBB_s"], a, b)
            loss, gR, gt = loss + ws * (l - 1.0), gR + ws * a, gt + ws * b
        if self.has_pharm:
            O, a, b = ch["pharm"]
            sim = self.similarity.lower()
            if sim == "tanimoto":
                l, a, b = tanimoto_chain(O, so["VAA_p"] + so["VBB_p"], a, b)
            else:
                s = SIGMA[sim]
                l, a, b = tversky_chain(O, s * so["VAA_p"] + (1 - s) * so["VBB_p"], a, b)
            loss, gR, gt = loss + wp * (l - 1.0), gR + wp * a, gt + wp * b
        return loss + 1.0, gR, gt

    # -- batched SE(3) loss + gradient (ag/_torch.py:602-772, 939-1017)
    def loss_grad_se3(self, se3, so):
        B = se3.shape[0]
        q_raw, t = se3[:, :4], se3[:, 4:]
        qn = torch.norm(q_raw, dim=1, keepdim=True)
        qh = q_raw / qn
        R = rot_from_unit_quat(qh)
        loss, gR, gt = self.loss_grad_Rt(R, t, so)
        gq = project_grad_R(gR, qh)
        proj = torch.eye(4, dtype=DT)[None] - qh.unsqueeze(-1) @ qh.unsqueeze(-2)
        gq_raw = (proj @ gq.unsqueeze(-1)).squeeze(-1) / qn
        return loss.mean(), torch.cat([gq_raw, gt], dim=1) / B, loss

    # -- final re-scoring through the scoring restatement (independent of the analytic path)
    def score_poses(self, se3):
        T = se3_matrix(se3)
        B = se3.shape[0]
        total = torch.zeros(B, dtype=DT)
        ws = (1.0 - self.color_weight) if self.mode == "shape_pharm" else 1.0
        wp = self.color_weight if self.mode == "shape_pharm" else 1.0
        if self.has_shape:
            fit_al = apply_se3(self.tfit.expand(B, -1, -1), T)
            ref = self.tref.expand(B, -1, -1)
            if self.mode == "esp":
                total = total + ws * tanimoto_esp(ref, fit_al, self.tref_esp.expand(B, -1),
                                                  self.tfit_esp.expand(B, -1), self.alpha, self.lam)
            else:
                total = total + ws * tanimoto_shape(ref, fit_al, self.alpha)
        if self.has_pharm:
            anc = apply_se3(self.tfit_anc.expand(B, -1, -1), T)
            vec = apply_so3(self.tfit_vec.expand(B, -1, -1), T)
            total = total + wp * pharm_similarity(self.tref_typ, self.tfit_typ, self.tref_anc, anc,
                                                  self.tref_vec, vec, self.similarity, **self.ph_kw)
        return total


# ------------------------------------------------------------------------------------------
# Scoring restatement (score/gaussian_overlap.py, electrostatic_scoring.py, pharmacophore_scoring.py)
# ------------------------------------------------------------------------------------------
def vab(c1, c2, alpha, w=None):
    """sum K exp(-alpha/2 r^2) [* w] over pairs, batched (B,N,3),(B,M,3).  gaussian_overlap.py:16-59."""
    r2 = ((c1.unsqueeze(2) - c2.unsqueeze(1)) ** 2).sum(-1)
    e = (math.pi ** 1.5) / ((2 * alpha) ** 1.5) * torch.exp(-(alpha / 2) * r2)
    if w is not None:
        e = e * w
    return e.sum(dim=(1, 2))


def tanimoto_shape(c1, c2, alpha):
    """gaussian_overlap.py:62-70."""
    a, b, ab = vab(c1, c1, alpha), vab(c2, c2, alpha), vab(c1, c2, alpha)
    return ab / (a + b - ab)


def tanimoto_esp(c1, c2, q1, q2, alpha, lam):
    """electrostatic_scoring.py:17-66 (charges (B,N),(B,M))."""
    w = lambda x, y: torch.exp(-((x.unsqueeze(2) - y.unsqueeze(1)) ** 2) / lam)
    a = vab(c1, c1, alpha, w(q1, q1))
    b = vab(c2, c2, alpha, w(q2, q2))
    ab = vab(c1, c2, alpha, w(q1, q2))
    return ab / (a + b - ab)


def _vab_cos(c1, c2, v1, v2, alpha, antiparallel):
    """gaussian_overlap.py:283-363."""
    v1 = F.normalize(v1, p=2, dim=-1)
    v2 = F.normalize(v2, p=2, dim=-1)
    cos = torch.bmm(v1, v2.transpose(1, 2))
    cos = cos.abs() if antiparallel else cos.clamp(0.0, 1.0)
    r2 = ((c1.unsqueeze(2) - c2.unsqueeze(1)) ** 2).sum(-1)
    e = (math.pi ** 1.5) / ((2 * alpha) ** 1.5) * ((cos + 2.0) / 3.0) * torch.exp(-(alpha / 2) * r2)
    return e.sum(dim=(1, 2))


def pharm_similarity(t1, t2, a1, a2, v1, v2, similarity="tanimoto", directionless=False,
                     extended_points=False, only_extended=False):
    """Per-type sum of (cosine-weighted) overlaps, then Tanimoto / Tversky.
    pharmacophore_scoring.py:123-145, 422-912; with ``extended_points`` the directional types
    take plain anchor overlaps (skipped with ``only_extended``) plus overlaps of the extended
    points anchor + unit vector (pharmacophore_scoring.py:284-370, vectors normalised as in
    the batched call, :562-563); directionless scoring turns extended points off (:517-523).
    t1 (N,), t2 (M,) shared by every pose; a2, v2 (B,M,3) posed; a1, v1 (N,3)."""
    if directionless:
        extended_points = only_extended = False
    B = a2.shape[0]
    a1b, v1b = a1.expand(B, -1, -1), v1.expand(B, -1, -1)
    VAB = torch.zeros(B, dtype=DT)
    VAA = torch.zeros(B, dtype=DT)
    VBB = torch.zeros(B, dtype=DT)
    present = set(int(x) for x in torch.cat([t1, t2]).unique())
    for ti in sorted(present):
        name = P_TYPES[ti]
        if name == "dummy":
            continue
        m1 = (t1 == ti)
        m2 = (t2 == ti)
        al = P_ALPHAS[name]
        c1, c2 = a1b[:, m1], a2[:, m2]
        if directionless or name in ("hydrophobe", "znbinder", "anion", "cation"):
            f = lambda x, y, vx, vy: vab(x, y, al)
        elif extended_points and name in ("acceptor", "donor", "halogen"):
            def f(x, y, vx, vy, al=al):
                ext = vab(x + F.normalize(vx, p=2, dim=-1), y + F.normalize(vy, p=2, dim=-1), al)
                return ext if only_extended else vab(x, y, al) + ext
        else:
            anti = name == "aromatic"
            f = lambda x, y, vx, vy: _vab_cos(x, y, vx, vy, al, anti)
        u1, u2 = v1b[:, m1], v2[:, m2]
        VAB = VAB + f(c1, c2, u1, u2)
        VAA = VAA + f(c1, c1, u1, u1)
        VBB = VBB + f(c2, c2, u2, u2)
    sim = similarity.lower()
    if sim == "tanimoto":
        return VAB / (VAA + VBB - VAB)
    s = SIGMA[sim]
    return torch.clamp_max(VAB / (s * VAA + (1 - s) * VBB), 1.0)


# ------------------------------------------------------------------------------------------
# The optimiser (alignment/_torch_analytical.py:23-478)
# ------------------------------------------------------------------------------------------
# ------------------------------------------------------------------------------------------
# ShaEP combination (vol_and_surf_esp): score/electrostatic_scoring.py:128-320 and the
# autograd optimiser alignment/_torch.py:804-1002
# ------------------------------------------------------------------------------------------
COULOMB_SCALING = 1e4 / (4 * 55.263 * np.pi)               # score/constants.py:16


def esp_comparison(points, atoms, charges, point_esp, radii, probe: float, lam_eff: float):
    """sum_i mask_i exp(-(pe_i - C sum_m q_m / d_im)^2 / lam_eff), batched: points (B,N,3),
    atoms (B,M,3), charges (M,), point_esp (N,), radii (M,) -> (B,).  mask_i = 1 iff every
    d_im >= radius_m + probe (electrostatic_scoring.py:128-227; lam_eff = LAM_SCALING * lam)."""
    d = torch.cdist(points, atoms)
    mask = (d >= radii[None, None, :] + probe).all(dim=2).to(d.dtype)
    phi = COULOMB_SCALING * (charges[None, None, :] / d).sum(dim=2)
    return (mask * torch.exp(-(point_esp[None, :] - phi) ** 2 / lam_eff)).sum(dim=1)


@dataclass
class ComboPair:
    """One pair under the ShaEP score (1-w) shape Tanimoto + w (esp_1 + esp_2) / (N_s + M_s).

    Shape centres (heavy atoms, or surface points when alpha != 0.81), surface points with their
    ESP, atoms including hydrogens with partial charges and vdW radii, for ref and fit.  ``lam`` is
    the raw lambda (LAM_SCALING is applied here).  The gradient is torch autograd of the score,
    as in the reference's optimiser."""
    ref_centers: np.ndarray
    fit_centers: np.ndarray
    ref_surf: np.ndarray
    fit_surf: np.ndarray
    ref_surf_esp: np.ndarray
    fit_surf_esp: np.ndarray
    ref_atoms: np.ndarray
    fit_atoms: np.ndarray
    ref_charges: np.ndarray
    fit_charges: np.ndarray
    ref_radii: np.ndarray
    fit_radii: np.ndarray
    alpha: float = 0.81
    lam: float = 0.001
    probe_radius: float = 1.0
    esp_weight: float = 0.5
    mode: str = "esp_combo"

    def __post_init__(self):
        for k in ("ref_centers", "fit_centers", "ref_surf", "fit_surf", "ref_surf_esp", "fit_surf_esp",
                  "ref_atoms", "fit_atoms", "ref_charges", "fit_charges", "ref_radii", "fit_radii"):
            setattr(self, "t" + k, torch.as_tensor(np.asarray(getattr(self, k), np.float64), dtype=DT))
        self.has_shape, self.has_pharm = True, False

    def self_overlaps(self):
        return {}

    def score_poses(self, se3):
        """Per-pose ShaEP score (B,) of poses se3 (B,7) (differentiable in se3)."""
        T = se3_matrix(se3)
        B = se3.shape[0]
        e = lambda x: x.expand(B, *x.shape)
        shape = tanimoto_shape(e(self.tref_centers), apply_se3(e(self.tfit_centers), T), self.alpha)
        lam_eff = LAM_SCALING * self.lam
        esp1 = esp_comparison(e(self.tref_surf), apply_se3(e(self.tfit_atoms), T), self.tfit_charges,
                              self.tref_surf_esp, self.tfit_radii, self.probe_radius, lam_eff)
        esp2 = esp_comparison(apply_se3(e(self.tfit_surf), T), e(self.tref_atoms), self.tref_charges,
                              self.tfit_surf_esp, self.tref_radii, self.probe_radius, lam_eff)
        nrm = float(len(self.ref_surf) + len(self.fit_surf))
        return self.esp_weight * (esp1 + esp2) / nrm + (1.0 - self.esp_weight) * shape

    @torch.enable_grad()
    def esp_value_grad(self, R, t):
        """esp_1 + esp_2 (B,) and its gradient w.r.t. R (B,3,3) and t (B,3) by autograd."""
        R = R.detach().clone().requires_grad_(True)
        t = t.detach().clone().requires_grad_(True)
        B = R.shape[0]
        e = lambda x: x.expand(B, *x.shape)
        lam_eff = LAM_SCALING * self.lam
        fa = e(self.tfit_atoms) @ R.transpose(1, 2) + t[:, None, :]
        fs = e(self.tfit_surf) @ R.transpose(1, 2) + t[:, None, :]
        v = esp_comparison(e(self.tref_surf), fa, self.tfit_charges, self.tref_surf_esp, self.tfit_radii,
                           self.probe_radius, lam_eff) + \
            esp_comparison(fs, e(self.tref_atoms), self.tref_charges, self.tfit_surf_esp, self.tref_radii,
                           self.probe_radius, lam_eff)
        gR, gt = torch.autograd.grad(v.sum(), (R, t))
        return v.detach(), gR, gt

    @torch.enable_grad()
    def loss_grad_se3(self, se3, so):
        x = se3.detach().clone().requires_grad_(True)
        per = 1.0 - self.score_poses(x)
        loss = per.mean()
        (g,) = torch.autograd.grad(loss, x)
        return loss.detach(), g, per.detach()


def optimize(pair: Pair, q0, t0, lr: float = 0.1, max_steps: int = 200, early_stop: bool = True,
             es_tol: float = 1e-5, es_patience: int = 10):
    """Multi-restart alignment with explicit seeds q0 (B,4), t0 (B,3).

    Returns dict(score (best), best_idx, scores (B,), se3 (B,7) final, transform (4,4), n_steps).
    """
    se3 = torch.cat([torch.as_tensor(np.asarray(q0, np.float64)),
                     torch.as_tensor(np.asarray(t0, np.float64))], dim=1).to(DT)
    se3.requires_grad_(True)
    so = pair.self_overlaps()
    opt = torch.optim.Adam([se3], lr=lr)
    # the analytic optimisers start the early-stop comparison at 1.0
    # (_torch_analytical.py:110, 283, 420), the ShaEP optimiser at +inf (alignment/_torch.py:922)
    last = math.inf if pair.mode == "esp_combo" else 1.0
    counter, n_steps = 0, max_steps
    for step in range(max_steps):
        with torch.no_grad():
            loss, grad, _ = pair.loss_grad_se3(se3, so)
        opt.zero_grad()
        se3.grad = grad
        opt.step()
        if early_stop:
            L = loss.item()
            counter = 0 if abs(L - last) > es_tol else counter + 1
            last = L
            if counter > es_patience:
                n_steps = step + 1
                break
    se3 = se3.detach()
    scores = pair.score_poses(se3)
    best = int(torch.argmax(scores))
    T = se3_matrix(se3)[best]
    return {"score": float(scores[best]), "best_idx": best, "scores": scores.numpy(),
            "se3": se3.numpy(), "transform": T.numpy(), "n_steps": n_steps}
