"""Drop-in pair containers with ``backend="hip"`` (rdkit-free).

``MoleculePairBatch.align_with_<mode>`` keeps the reference's method names, keyword arguments,
defaults, in-place result attributes and return contract
(shepherd_score/container/_batch.py:274-1515; attributes accel/_modes.py:18-26):
  * results land on each pair as ``transform_<mode>`` (4x4 float32 torch tensor, fit -> ref,
    applied as ``fit @ R.T + t``) and ``sim_aligned_<mode>`` (Python float);
  * the call returns ``(scores np.float64 (K,), aligned list or [None] * K)``; pharm returns
    ``(scores, anchors list, vectors list)``.
The backend dispatch mirrors ``_run_fast_or_fallthrough`` (_batch.py:127-151): ``None``, ``"hip"``
and the accelerated names ``"triton"``, ``"cuda"``, ``"gpu"`` run the gfx950 kernels; names of
backends this package does not ship (``"numba"``, ``"cpu"``, ``"jax"``) raise ValueError.

Semantics of the HIP backend: ``num_repeats`` restarts from the reference's multi-start seeds
(identity, four principal-axis alignments, Fibonacci rotations; alignment/_torch.py:253-326),
``lr`` and ``max_num_steps`` are honoured, torch-style Adam, per-pair early stop
(alignment/_torch_analytical.py:23-478), FP64 re-scoring and best-restart selection.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .align import OptimSpec, ScoreSpec, align, pose_score, self_overlaps
from .constants import LAM_SCALING
from .modes import HIP_MODES, MODE_ATTRS, MODE_SEEDS, MODE_STEPS, canonical
from .packing import CloudSet, PairBatch, PharmSet
from .seeds import device_seeds, initial_seeds, pad

HIP_BACKENDS = ("hip", "rocm", "triton", "cuda", "gpu")


def resolve_backend(backend: Optional[str]) -> str:
    if backend is None or backend in HIP_BACKENDS:
        return "hip"
    raise ValueError(f"unknown backend {backend!r}; fast_shepherd_score_amd runs backend='hip' "
                     f"(aliases {HIP_BACKENDS[1:]})")


class Molecule:
    """Array-only molecule: heavy-atom positions and optional surface / ESP / pharmacophores.

    Mirrors the attributes the reference's aligners read from a ``Molecule``
    (container/_core.py; screen.py:650-668): ``atom_pos``, ``surf_pos``, ``surf_esp``,
    ``pharm_types``, ``pharm_ancs``, ``pharm_vecs``, per-heavy-atom ``atom_charges``, and for
    vol_and_surf_esp the atoms including hydrogens ``centers_w_H`` with their
    ``partial_charges`` and vdW ``radii`` (the reference reads them from the RDKit conformer).
    """

    def __init__(self, atom_pos, surf_pos=None, surf_esp=None, pharm_types=None,
                 pharm_ancs=None, pharm_vecs=None, atom_charges=None, centers_w_H=None,
                 partial_charges=None, radii=None):
        f = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, np.float32))
        self.atom_pos = f(atom_pos)
        self.surf_pos = f(surf_pos)
        self.surf_esp = f(surf_esp)
        self.pharm_types = None if pharm_types is None else np.asarray(pharm_types, np.int64)
        self.pharm_ancs = f(pharm_ancs)
        self.pharm_vecs = f(pharm_vecs)
        self.atom_charges = f(atom_charges)
        self.centers_w_H = f(centers_w_H)
        self.partial_charges = f(partial_charges)
        self.radii = f(radii)

    def centered(self) -> "Molecule":
        c = self.atom_pos.mean(0)
        sub = lambda a: None if a is None else a - c
        return Molecule(self.atom_pos - c, sub(self.surf_pos), self.surf_esp, self.pharm_types,
                        sub(self.pharm_ancs), self.pharm_vecs, self.atom_charges,
                        sub(self.centers_w_H), self.partial_charges, self.radii)


class MoleculePair:
    """A (ref, fit) pair; the fit molecule is moved onto the ref by the alignment."""

    def __init__(self, ref_mol: Molecule, fit_mol: Molecule, num_surf_points: Optional[int] = None,
                 do_center: bool = False, device=None):
        self.ref_molec = ref_mol.centered() if do_center else ref_mol
        self.fit_molec = fit_mol.centered() if do_center else fit_mol
        self.num_surf_points = num_surf_points
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        for tf, sc in list(MODE_ATTRS.values()) + list(_ATTRS_WITH_H.values()):
            setattr(self, tf, None)
            setattr(self, sc, None)

    # -- single-pair API (container/_core.py:612-1357) ---------------------------------------
    # The same persistent launch as the batch, over one pair.  Like the reference's per-pair
    # methods it stores the transform as a numpy (4, 4) array and the score as a numpy scalar,
    # and returns the aligned fit cloud.  ``use_jax`` / ``use_analytical`` / ``verbose`` are
    # accepted and have no effect: the HIP kernels are the implementation.
    def _align_one(self, mode, cloud, num_repeats, max_num_steps, no_H=True, **kw):
        if num_repeats is None:
            num_repeats = MODE_SEEDS[mode]
        if max_num_steps is None:
            max_num_steps = MODE_STEPS[mode]
        align_pairs([self], mode, num_repeats=num_repeats, max_num_steps=max_num_steps, no_H=no_H, **kw)
        tf_attr, sc_attr = result_attrs(mode, no_H)
        T = getattr(self, tf_attr).numpy()
        setattr(self, tf_attr, T)
        setattr(self, sc_attr, np.asarray(getattr(self, sc_attr)))
        return self._moved(cloud, T)

    def _moved(self, cloud, T):
        X = np.asarray(getattr(self.fit_molec, cloud), np.float64)
        Td = np.asarray(T, np.float64)
        return (X @ Td[:3, :3].T + Td[:3, 3]).astype(np.float32)

    def align_with_vol(self, no_H: bool = True, num_repeats: int = None, trans_init: bool = False,
                       lr: float = 0.1, max_num_steps: int = None, use_jax: bool = False,
                       use_analytical: bool = True, verbose: bool = False) -> np.ndarray:
        return self._align_one("vol", "atom_pos" if no_H else "centers_w_H", num_repeats,
                               max_num_steps, no_H=no_H, trans_init=trans_init, lr=lr)

    def align_with_vol_esp(self, lam: float = 0.1, no_H: bool = True, num_repeats: int = None,
                           trans_init: bool = False, lr: float = 0.1, max_num_steps: int = None,
                           use_jax: bool = False, use_analytical: bool = True,
                           verbose: bool = False) -> np.ndarray:
        """``lam`` is raw here, as in the reference (no LAM_SCALING)."""
        return self._align_one("vol_esp", "atom_pos" if no_H else "centers_w_H", num_repeats,
                               max_num_steps, no_H=no_H, trans_init=trans_init, lr=lr, lam=lam)

    def align_with_surf(self, alpha: float, num_repeats: int = None, trans_init: bool = False,
                        lr: float = 0.1, max_num_steps: int = None, use_jax: bool = False,
                        use_analytical: bool = True, verbose: bool = False) -> np.ndarray:
        return self._align_one("surf", "surf_pos", num_repeats, max_num_steps,
                               trans_init=trans_init, lr=lr, alpha=alpha)

    def align_with_surf_esp(self, alpha: float, lam: float = 0.3, num_repeats: int = None,
                            trans_init: bool = False, lr: float = 0.1, max_num_steps: int = None,
                            use_jax: bool = False, use_analytical: bool = True,
                            verbose: bool = False) -> np.ndarray:
        """``lam`` is scaled by LAM_SCALING, as in the reference."""
        return self._align_one("surf_esp", "surf_pos", num_repeats, max_num_steps,
                               trans_init=trans_init, lr=lr, alpha=alpha, lam=LAM_SCALING * lam)

    def align_with_vol_and_surf_esp(self, alpha: float, lam: float = 0.001,
                                    probe_radius: float = 1.0, esp_weight: float = 0.5,
                                    num_repeats: int = None, trans_init: bool = False,
                                    lr: float = 0.1, max_num_steps: int = None,
                                    use_jax: bool = False, verbose: bool = False) -> np.ndarray:
        return self._align_one("vol_and_surf_esp", "surf_pos", num_repeats, max_num_steps,
                               trans_init=trans_init, lr=lr, alpha=alpha, lam=lam,
                               probe_radius=probe_radius, esp_weight=esp_weight)

    align_with_esp = align_with_surf_esp                  # legacy aliases (_core.py:1138-1139)
    align_with_esp_combo = align_with_vol_and_surf_esp

    def align_with_vol_color(self, color_weight: float = 0.5, alpha: float = 0.81,
                             similarity: str = "tanimoto", directionless: bool = True,
                             extended_points: bool = False, only_extended: bool = False,
                             num_repeats: int = None, trans_init: bool = False, lr: float = 0.1,
                             max_num_steps: int = None, verbose: bool = False) -> np.ndarray:
        return self._align_one("vol_color", "atom_pos", num_repeats, max_num_steps,
                               trans_init=trans_init, lr=lr, alpha=alpha, color_weight=color_weight,
                               similarity=similarity, directionless=directionless,
                               extended_points=extended_points, only_extended=only_extended)

    def align_with_pharm(self, similarity: str = "tanimoto", extended_points: bool = False,
                         only_extended: bool = False, num_repeats: int = None,
                         trans_init: bool = False, lr: float = 0.1, max_num_steps: int = None,
                         use_jax: bool = False, verbose: bool = False, use_vectorized: bool = True,
                         use_analytical: bool = True):
        """Returns (aligned anchors, aligned vectors), as the reference does."""
        anchors = self._align_one("pharm", "pharm_ancs", num_repeats, max_num_steps,
                                  trans_init=trans_init, lr=lr, similarity=similarity,
                                  extended_points=extended_points, only_extended=only_extended)
        R = np.asarray(self.transform_pharm, np.float64)[:3, :3]
        vectors = (np.asarray(self.fit_molec.pharm_vecs, np.float64) @ R.T).astype(np.float32)
        return anchors, vectors

    # -- scores of the current alignment (_core.py:1359-1560), FP64 on the device ----------
    # ``use`` is validated as the reference does; every value runs the HIP re-scoring kernel.
    # Pharmacophore vectors are normalised, as in the reference's batched scoring
    # (pharmacophore_scoring.py:562-563); unit vectors score the same either way.
    def _score_now(self, mode, use, **kw):
        use = use.lower()
        accepted = ("jax", "jnp", "torch", "pytorch", "np", "numpy")
        if use not in accepted:
            raise ValueError(f"`use` must be in {accepted}. Instead {use} was passed.")
        return score_pairs([self], mode, **kw)

    def score_with_surf(self, alpha: float, use: str = "np") -> np.ndarray:
        return self._score_now("surf", use, alpha=alpha)

    def score_with_esp(self, alpha: float, lam: float = 0.3, use: str = "np") -> np.ndarray:
        return self._score_now("surf_esp", use, alpha=alpha, lam=LAM_SCALING * lam)

    def score_with_pharm(self, similarity: str = "tanimoto", extended_points: bool = False,
                         only_extended: bool = False, use: str = "np") -> np.ndarray:
        return self._score_now("pharm", use, similarity=similarity, extended_points=extended_points,
                               only_extended=only_extended)


# vol / vol_esp with hydrogens (no_H=False) write the attributes without the _noH suffix
_ATTRS_WITH_H = {"vol": ("transform_vol", "sim_aligned_vol"),
                 "vol_esp": ("transform_vol_esp", "sim_aligned_vol_esp")}


def result_attrs(mode: str, no_H: bool = True):
    """(transform attribute, score attribute) a mode writes (_modes.py:18-26, _core.py:705-710)."""
    if not no_H and mode in _ATTRS_WITH_H:
        return _ATTRS_WITH_H[mode]
    return MODE_ATTRS[mode]


def _mode_inputs(pairs: Sequence[MoleculePair], mode: str, no_H: bool = True):
    """Host arrays (ref clouds, fit clouds, ref w, fit w, pharm dict) a mode consumes.
    ``no_H=False`` (vol, vol_esp) takes the atoms with hydrogens ``centers_w_H`` and their
    ``partial_charges``, as ``Molecule.get_positions(no_H=False)`` does (_core.py:332-350)."""
    R = [p.ref_molec for p in pairs]
    F = [p.fit_molec for p in pairs]
    need = lambda ms, a: [getattr(m, a) for m in ms]
    def check(vals, what):
        if any(v is None for v in vals):
            raise ValueError(f"mode needs {what} on every molecule")
        return vals
    ph = None
    rw = fw = None
    if not no_H and mode in ("vol", "vol_esp"):
        rc, fc = check(need(R, "centers_w_H"), "centers_w_H"), check(need(F, "centers_w_H"), "centers_w_H")
        if mode == "vol_esp":
            rw = check(need(R, "partial_charges"), "partial_charges")
            fw = check(need(F, "partial_charges"), "partial_charges")
        return rc, fc, rw, fw, None
    if mode in ("vol", "vol_color"):
        rc, fc = check(need(R, "atom_pos"), "atom_pos"), check(need(F, "atom_pos"), "atom_pos")
    elif mode == "vol_esp":
        rc, fc = check(need(R, "atom_pos"), "atom_pos"), check(need(F, "atom_pos"), "atom_pos")
        rw, fw = check(need(R, "atom_charges"), "atom_charges"), check(need(F, "atom_charges"), "atom_charges")
    elif mode == "surf":
        rc, fc = check(need(R, "surf_pos"), "surf_pos"), check(need(F, "surf_pos"), "surf_pos")
    elif mode == "surf_esp":
        rc, fc = check(need(R, "surf_pos"), "surf_pos"), check(need(F, "surf_pos"), "surf_pos")
        rw, fw = check(need(R, "surf_esp"), "surf_esp"), check(need(F, "surf_esp"), "surf_esp")
    elif mode == "pharm":
        rc, fc = None, None
    elif mode == "vol_and_surf_esp":     # shape centres are chosen by alpha in align_pairs
        rc, fc = check(need(R, "atom_pos"), "atom_pos"), check(need(F, "atom_pos"), "atom_pos")
    else:
        raise ValueError(mode)
    if mode in ("pharm", "vol_color"):
        ph = {}
        for side, ms in (("ref", R), ("fit", F)):
            ph[f"{side}_types"] = check(need(ms, "pharm_types"), "pharm_types")
            ph[f"{side}_anchors"] = check(need(ms, "pharm_ancs"), "pharm_ancs")
            ph[f"{side}_vectors"] = check(need(ms, "pharm_vecs"), "pharm_vecs")
    return rc, fc, rw, fw, ph


def _shaep_sets(pairs: Sequence[MoleculePair], dev):
    """ShaEP channel of both sides: surfaces with ESP, atoms with hydrogens, charges, radii."""
    from .packing import ShaepSet
    out = []
    for side in ("ref_molec", "fit_molec"):
        ms = [getattr(p, side) for p in pairs]
        for a in ("surf_pos", "surf_esp", "centers_w_H", "partial_charges", "radii"):
            if any(getattr(m, a) is None for m in ms):
                raise ValueError(f"vol_and_surf_esp needs {a} on every molecule")
        out.append(ShaepSet.from_arrays([m.surf_pos for m in ms], [m.surf_esp for m in ms],
                                        [m.centers_w_H for m in ms], [m.partial_charges for m in ms],
                                        [m.radii for m in ms], device=dev))
    return out


def _trans_init_align(pairs, mode, kw):
    """trans_init=True: the seeds of _initialize_se3_params_with_translations (5 + 10 per
    centre; centres = the ref's atoms, or its pharmacophore anchors in pharm mode; _core.py:671,
    1094, 1228, 1325) built per pair on the host, then one persistent launch per group of pairs
    with the same restart count (the optimiser's mean loss and gradient scale are per pair)."""
    from .seeds import translation_seeds
    groups = {}
    for p in pairs:
        r, f = p.ref_molec, p.fit_molec
        if mode == "pharm":
            sref, sfit, cen = r.pharm_ancs, f.pharm_ancs, r.pharm_ancs
        else:
            if mode == "vol_and_surf_esp":        # seeded from the surface points (_torch.py:906)
                sref, sfit = r.surf_pos, f.surf_pos
            else:
                rc, fc, _, _, _ = _mode_inputs([p], mode, kw.get("no_H", True))
                sref, sfit = rc[0], fc[0]
            cen = r.atom_pos
        if sref is None or sfit is None or cen is None:
            raise ValueError(f"trans_init needs the {mode} clouds and the ref's translation centres")
        q, t = translation_seeds(sref, sfit, cen)
        if kw.get("early_stop", True) and len(q) > 1024:
            raise ValueError(f"trans_init gives {len(q)} restarts for a pair; per-pair early stop keeps a "
                             "pair's restarts in one workgroup (at most 1024, i.e. 101 centres)")
        groups.setdefault(len(q), []).append((p, q, t))
    res = None
    for items in groups.values():
        seeds = (np.stack([it[1] for it in items]), np.stack([it[2] for it in items]))
        res = align_pairs([it[0] for it in items], mode, seeds=seeds, **kw)
    return res


def align_pairs(pairs: Sequence[MoleculePair], mode: str, *, alpha=0.81, lam=None,
                color_weight=0.5, similarity="tanimoto", directionless=False,
                extended_points=False, only_extended=False, num_repeats: int = 50, lr: float = 0.1, max_num_steps: int = 200,
                early_stop: bool = True, seeds=None, device=None, probe_radius: float = 1.0,
                esp_weight: float = 0.5, trans_init: bool = False, no_H: bool = True, **tiling):
    """Align every pair in one persistent launch; write results in place; return the result.

    ``seeds``: None computes the reference's multi-start seeds on the device (fss_seed_poses);
    "host" computes the same rule on the host in FP64 (seeds.initial_seeds); a (q0, t0) pair of
    (n_pairs, S, 4) / (n_pairs, S, 3) arrays or tensors is used as given.  ``trans_init``: the
    reference's translation-centre seeds instead (num_repeats is then unused); pairs with
    different restart counts run as separate launches.  ``no_H=False`` (vol, vol_esp) aligns
    the atoms with hydrogens and writes ``transform_vol[_esp]``.  ``extended_points`` /
    ``only_extended``: the pharmacophore channel scores acceptors, donors and halogens by
    extended-point overlaps (pharmacophore_scoring.py:284-370); off when ``directionless``.
    """
    mode = canonical(mode)
    if mode not in HIP_MODES:
        raise NotImplementedError(f"mode {mode!r} is not implemented by the HIP backend")
    if not pairs:
        return None
    if trans_init:
        if seeds is not None:
            raise ValueError("trans_init builds its own seeds: do not pass seeds")
        kw = dict(alpha=alpha, lam=lam, color_weight=color_weight, similarity=similarity,
                  directionless=directionless, extended_points=extended_points,
                  only_extended=only_extended, lr=lr, max_num_steps=max_num_steps,
                  early_stop=early_stop, device=device, probe_radius=probe_radius,
                  esp_weight=esp_weight, no_H=no_H, **tiling)
        return _trans_init_align(pairs, canonical(mode), kw)
    pb, spec, (rc, fc, ph), dev = _pair_batch(pairs, mode, alpha, lam, color_weight, similarity,
                                              directionless, extended_points, only_extended,
                                              probe_radius, esp_weight, no_H, device)
    combo = mode == "vol_and_surf_esp"
    # Seeds come from the clouds the objective moves: pharm anchors for pharm, the surface
    # points for vol_and_surf_esp (alignment/_torch.py:906), else the shape clouds.
    if seeds is None:
        q0, t0 = device_seeds(pb, num_repeats, use_pharm=mode == "pharm", use_surface=combo)
    elif isinstance(seeds, str) and seeds == "host":
        sref = ph["ref_anchors"] if mode == "pharm" else ([p.ref_molec.surf_pos for p in pairs] if combo else rc)
        sfit = ph["fit_anchors"] if mode == "pharm" else ([p.fit_molec.surf_pos for p in pairs] if combo else fc)
        (R, rn), (Fp, fn) = pad(sref), pad(sfit)
        q0, t0 = initial_seeds(R, rn, Fp, fn, num_repeats)
    else:
        q0, t0 = seeds
    q0 = torch.as_tensor(np.asarray(q0, np.float32) if not torch.is_tensor(q0) else q0,
                         dtype=torch.float32).to(dev)
    t0 = torch.as_tensor(np.asarray(t0, np.float32) if not torch.is_tensor(t0) else t0,
                         dtype=torch.float32).to(dev)
    opt = OptimSpec(max_steps=int(max_num_steps), lr=float(lr), early_stop=bool(early_stop))
    res = align(pb, spec, q0, t0, opt, **tiling)
    scores = res.score.cpu().numpy()
    tfs = res.transform.cpu()
    tf_attr, sc_attr = result_attrs(mode, no_H)
    for i, p in enumerate(pairs):
        setattr(p, tf_attr, tfs[i].clone())
        setattr(p, sc_attr, float(scores[i]))
    return res


def _pair_batch(pairs, mode, alpha, lam, color_weight, similarity, directionless, extended_points,
                only_extended, probe_radius, esp_weight, no_H, device):
    """The device PairBatch and ScoreSpec of a mode (with its default lam), and the host clouds."""
    dev = torch.device(device) if device is not None else pairs[0].device
    rc, fc, rw, fw, ph = _mode_inputs(pairs, mode, no_H)
    kind = {"vol": _lib.MODE_SHAPE, "surf": _lib.MODE_SHAPE, "vol_esp": _lib.MODE_SHAPE_ESP,
            "surf_esp": _lib.MODE_SHAPE_ESP, "pharm": _lib.MODE_PHARM,
            "vol_color": _lib.MODE_SHAPE_PHARM, "vol_and_surf_esp": _lib.MODE_ESP_COMBO}[mode]
    combo = mode == "vol_and_surf_esp"
    if combo:
        # volumetric shape term at alpha 0.81, surface shape term otherwise (_core.py:1079-1080);
        # lam is the comparison's raw lambda (default 0.001), LAM_SCALING applied here
        if alpha != 0.81:
            rc = [p.ref_molec.surf_pos for p in pairs]
            fc = [p.fit_molec.surf_pos for p in pairs]
        lam = LAM_SCALING * (0.001 if lam is None else lam)
    elif lam is None:
        lam = 0.3 * LAM_SCALING if mode == "surf_esp" else 0.1
    spec = ScoreSpec(mode=kind, alpha=alpha, lam=lam, color_weight=color_weight,
                     similarity=similarity, directionless=directionless,
                     extended_points=bool(extended_points) and not directionless,
                     only_extended=bool(only_extended) and not directionless,
                     esp_weight=esp_weight, probe_radius=probe_radius)
    ref = CloudSet.from_arrays(rc, rw, device=dev) if rc is not None else None
    fit = CloudSet.from_arrays(fc, fw, device=dev) if fc is not None else None
    rph = fph = None
    if ph is not None:
        rph = PharmSet.from_arrays(ph["ref_types"], ph["ref_anchors"], ph["ref_vectors"], device=dev)
        fph = PharmSet.from_arrays(ph["fit_types"], ph["fit_anchors"], ph["fit_vectors"], device=dev)
    rse = fse = None
    if combo:
        rse, fse = _shaep_sets(pairs, dev)
    pb = PairBatch(ref, fit, rph, fph, ref_se=rse, fit_se=fse)
    return pb, spec, (rc, fc, ph), dev


def score_pairs(pairs: Sequence[MoleculePair], mode: str, *, alpha=0.81, lam=None,
                color_weight=0.5, similarity="tanimoto", directionless=False,
                extended_points=False, only_extended=False, probe_radius: float = 1.0,
                esp_weight: float = 0.5, no_H: bool = True, device=None) -> np.ndarray:
    """FP64 similarity of every pair at its current pose (no alignment), (n_pairs,) float64:
    the re-scoring kernel at the identity pose (gaussian_overlap.py:73-106,
    electrostatic_scoring.py:69-125, pharmacophore_scoring.py:422-912)."""
    mode = canonical(mode)
    if mode not in HIP_MODES:
        raise NotImplementedError(f"mode {mode!r} is not implemented by the HIP backend")
    if not pairs:
        return np.zeros(0)
    pb, spec, _, dev = _pair_batch(pairs, mode, alpha, lam, color_weight, similarity, directionless,
                                   extended_points, only_extended, probe_radius, esp_weight, no_H,
                                   device)
    n = len(pairs)
    q = torch.zeros((n, 4), dtype=torch.float32)
    q[:, 0] = 1.0
    so = self_overlaps(pb, spec)
    return pose_score(pb, spec, torch.arange(n, dtype=torch.int32), q, torch.zeros((n, 3)), so).cpu().numpy()


class MoleculePairBatch:
    """Batch of MoleculePair objects aligned together on the GPU."""

    _TRITON_BACKENDS = ("triton", "cuda", "gpu")
    _HIP_BACKENDS = HIP_BACKENDS

    def __init__(self, pairs: List[MoleculePair]):
        self.pairs = pairs

    # -- shared plumbing -------------------------------------------------------------------
    def _run(self, mode, backend, return_aligned, fit_cloud, num_repeats, max_num_steps, **kw):
        resolve_backend(backend)
        if num_repeats is None:
            num_repeats = MODE_SEEDS[mode]
        if max_num_steps is None:
            max_num_steps = MODE_STEPS[mode]
        align_pairs(self.pairs, mode, num_repeats=num_repeats, max_num_steps=max_num_steps, **kw)
        tf_attr, sc_attr = result_attrs(mode, kw.get("no_H", True))
        scores = np.array([float(getattr(p, sc_attr)) for p in self.pairs])
        if not return_aligned:
            return scores, [None] * len(self.pairs)
        aligned = []
        for p in self.pairs:
            T = getattr(p, tf_attr).numpy().astype(np.float64)
            X = np.asarray(getattr(p.fit_molec, fit_cloud), np.float64)
            aligned.append((X @ T[:3, :3].T + T[:3, 3]).astype(np.float32))
        return scores, aligned

    def align_with_vol_and_surf_esp(self, alpha: float, lam: float = 0.001, probe_radius: float = 1.0,
                                    esp_weight: float = 0.5, num_repeats: int = None,
                                    trans_init: bool = False, lr: float = 0.1,
                                    max_num_steps: int = None, verbose: bool = False,
                                    backend: Optional[str] = None, return_aligned: bool = False,
                                    num_workers: int = 1, use_shmap: bool = False, **hip_kw):
        """ShaEP score: (1 - esp_weight) shape Tanimoto (heavy atoms at alpha 0.81, surface
        points otherwise) + esp_weight surface ESP comparison (_batch.py:1030-1102)."""
        return self._run("vol_and_surf_esp", backend, return_aligned, "surf_pos", num_repeats,
                         max_num_steps, trans_init=trans_init, lr=lr, alpha=alpha, lam=lam,
                         probe_radius=probe_radius, esp_weight=esp_weight, **self._hip_kw(hip_kw))

    align_with_esp_combo = align_with_vol_and_surf_esp   # legacy alias (_batch.py:1102)

    @staticmethod
    def _hip_kw(kw):
        keep = ("early_stop", "seeds", "lanes_per_pose", "restarts_per_wg", "device")
        return {k: v for k, v in kw.items() if k in keep}

    # -- public API (reference signatures) --------------------------------------------------
    def align_with_vol(self, no_H: bool = True, num_repeats: int = None, trans_init: bool = False,
                       lr: float = 0.1, max_num_steps: int = None, num_workers: int = 1,
                       use_shmap: bool = True, num_buckets: int = 1, verbose: bool = False,
                       backend: Optional[str] = None, alpha: float = 0.81,
                       return_aligned: bool = False, **hip_kw):
        return self._run("vol", backend, return_aligned, "atom_pos" if no_H else "centers_w_H",
                         num_repeats, max_num_steps, trans_init=trans_init, lr=lr, alpha=alpha,
                         no_H=no_H, **self._hip_kw(hip_kw))

    def align_with_vol_esp(self, lam: float = 0.1, no_H: bool = True, num_repeats: int = None,
                           trans_init: bool = False, lr: float = 0.1, max_num_steps: int = None,
                           num_workers: int = 1, use_shmap: bool = True, num_buckets: int = 1,
                           verbose: bool = False, backend: Optional[str] = None,
                           alpha: float = 0.81, return_aligned: bool = False, **hip_kw):
        return self._run("vol_esp", backend, return_aligned, "atom_pos" if no_H else "centers_w_H",
                         num_repeats, max_num_steps, trans_init=trans_init, lr=lr, alpha=alpha,
                         lam=lam, no_H=no_H, **self._hip_kw(hip_kw))

    def align_with_surf(self, alpha: float, num_repeats: int = None, trans_init: bool = False,
                        lr: float = 0.1, max_num_steps: int = None, use_jax: bool = True,
                        use_analytical: bool = True, num_workers: int = 1,
                        use_shmap: bool = False, verbose: bool = False,
                        backend: Optional[str] = None, return_aligned: bool = False, **hip_kw):
        return self._run("surf", backend, return_aligned, "surf_pos", num_repeats, max_num_steps,
                         trans_init=trans_init, lr=lr, alpha=alpha, **self._hip_kw(hip_kw))

    def align_with_surf_esp(self, alpha: float, lam: float = 0.3, num_repeats: int = None,
                            trans_init: bool = False, lr: float = 0.1, max_num_steps: int = None,
                            use_jax: bool = True, use_analytical: bool = True,
                            num_workers: int = 1, use_shmap: bool = False, verbose: bool = False,
                            backend: Optional[str] = None, return_aligned: bool = False, **hip_kw):
        # lam is scaled by LAM_SCALING internally, as in the reference (_batch.py:944-948).
        return self._run("surf_esp", backend, return_aligned, "surf_pos", num_repeats,
                         max_num_steps, trans_init=trans_init, lr=lr, alpha=alpha,
                         lam=LAM_SCALING * lam, **self._hip_kw(hip_kw))

    align_with_esp = align_with_surf_esp   # legacy alias (accel/_modes.py:35)

    def align_with_pharm(self, similarity: str = "tanimoto", extended_points: bool = False,
                         only_extended: bool = False, num_repeats: int = None,
                         trans_init: bool = False, lr: float = 0.1, max_num_steps: int = None,
                         num_workers: int = 1, use_shmap: bool = True, num_buckets: int = 1,
                         verbose: bool = False, backend: Optional[str] = None,
                         return_aligned: bool = False, **hip_kw):
        scores, _ = self._run("pharm", backend, False, None, num_repeats, max_num_steps,
                              trans_init=trans_init, lr=lr, similarity=similarity,
                              extended_points=extended_points, only_extended=only_extended,
                              **self._hip_kw(hip_kw))
        n = len(self.pairs)
        if not return_aligned:
            return scores, [None] * n, [None] * n
        anchors, vectors = [], []
        for p in self.pairs:
            T = p.transform_pharm.numpy().astype(np.float64)
            anchors.append((np.asarray(p.fit_molec.pharm_ancs, np.float64) @ T[:3, :3].T + T[:3, 3]).astype(np.float32))
            vectors.append((np.asarray(p.fit_molec.pharm_vecs, np.float64) @ T[:3, :3].T).astype(np.float32))
        return scores, anchors, vectors

    def align_with_vol_color(self, color_weight: float = 0.5, alpha: float = 0.81,
                             num_repeats: int = None, trans_init: bool = False, lr: float = 0.1,
                             max_num_steps: int = None, verbose: bool = False,
                             backend: Optional[str] = None, return_aligned: bool = False,
                             directionless: bool = True, similarity: str = "tanimoto",
                             extended_points: bool = False, only_extended: bool = False,
                             **hip_kw):
        """Shape + pharmacophore colour.  ``directionless=False`` keeps the directional
        pharmacophore channel, cosine-weighted or with ``extended_points``
        (alignment/_torch.py:1376-1548, container/_core.py:1142-1236)."""
        return self._run("vol_color", backend, return_aligned, "atom_pos", num_repeats,
                         max_num_steps, trans_init=trans_init, lr=lr, alpha=alpha,
                         color_weight=color_weight, directionless=directionless,
                         similarity=similarity, extended_points=extended_points,
                         only_extended=only_extended, **self._hip_kw(hip_kw))
