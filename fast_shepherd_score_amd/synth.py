"""Seeded synthetic molecules for the benchmark workloads (numpy, deterministic).

* atoms: a 3D random walk from the origin with 1.5 A steps in uniform random directions; a
  direction is redrawn (up to 100 tries) until the new atom is >= 1.2 A from every earlier
  atom; the walk is then centred on its mean.
* surface: points drawn uniformly on the exposed part of the union of spheres of radius
  1.2 * 1.7 A around the atoms (a draw inside any other sphere is rejected).
* ESP values: N(0, 0.1) clipped to +-0.3.
* pharmacophores: types drawn uniformly from the 8 real types (0..7), anchor = a random atom
  + N(0, 0.5 A), unit direction vectors for acceptor / donor / aromatic / halogen and zero
  vectors for the others.
"""
from __future__ import annotations

import numpy as np

STEP = 1.5
MIN_SEP = 1.2
VDW = 1.7
SURF_SCALE = 1.2
DIRECTIONAL_TYPE_IDS = (0, 1, 2, 4)   # acceptor, donor, aromatic, halogen


def _unit(rng, shape):
    v = rng.normal(size=shape + (3,))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_walks(rng: np.random.Generator, counts) -> list:
    """Random-walk heavy-atom clouds, one per entry of ``counts`` (vectorised over molecules)."""
    counts = np.asarray(counts, dtype=np.int64)
    n = len(counts)
    L = int(counts.max()) if n else 0
    X = np.zeros((n, max(L, 1), 3))
    for k in range(1, L):
        live = np.nonzero(counts > k)[0]
        prev = X[live, k - 1]
        pos = prev + STEP * _unit(rng, (len(live),))
        todo = np.ones(len(live), dtype=bool)
        for _ in range(100):
            d = np.linalg.norm(X[live, :k] - pos[:, None, :], axis=-1)
            ok = (d >= MIN_SEP).all(axis=1)
            todo &= ~ok
            if not todo.any():
                break
            idx = np.nonzero(todo)[0]
            pos[idx] = prev[idx] + STEP * _unit(rng, (len(idx),))
        X[live, k] = pos
    out = []
    for i in range(n):
        a = X[i, :counts[i]]
        out.append((a - a.mean(axis=0)).astype(np.float32))
    return out


def surface_points(rng: np.random.Generator, atoms: np.ndarray, n_points: int) -> np.ndarray:
    """n_points uniform points on the exposed union of 1.2*vdW spheres around ``atoms``."""
    r = SURF_SCALE * VDW
    atoms = np.asarray(atoms, np.float64)
    got = []
    need = n_points
    while need > 0:
        m = max(2 * need, 64)
        which = rng.integers(0, len(atoms), size=m)
        pts = atoms[which] + r * _unit(rng, (m,))
        d = np.linalg.norm(pts[:, None, :] - atoms[None, :, :], axis=-1)
        d[np.arange(m), which] = np.inf
        keep = pts[(d >= r - 1e-9).all(axis=1)]
        got.append(keep[:need])
        need -= len(got[-1])
    return np.concatenate(got, axis=0).astype(np.float32)


def esp_values(rng: np.random.Generator, n: int) -> np.ndarray:
    return np.clip(rng.normal(0.0, 0.1, size=n), -0.3, 0.3).astype(np.float32)


def pharmacophores(rng: np.random.Generator, atoms: np.ndarray, n_ph: int):
    """(types int64 (n_ph,), anchors (n_ph, 3), vectors (n_ph, 3))."""
    types = rng.integers(0, 8, size=n_ph)
    anchors = atoms[rng.integers(0, len(atoms), size=n_ph)] + rng.normal(0.0, 0.5, size=(n_ph, 3))
    vecs = _unit(rng, (n_ph,))
    vecs[~np.isin(types, DIRECTIONAL_TYPE_IDS)] = 0.0
    return types.astype(np.int64), anchors.astype(np.float32), vecs.astype(np.float32)


H_BOND = 1.09
VDW_H = 1.1
SHAEP_PROBE = 1.2
COULOMB = 1e4 / (4 * 55.263 * np.pi)


def shaep_molecule(rng: np.random.Generator, n_heavy: int, n_surf: int) -> dict:
    """A molecule for the ShaEP (vol_and_surf_esp) score: random-walk heavy atoms (vdW 1.7 A),
    one or two hydrogens (1.09 A bonds, vdW 1.1 A) on each, partial charges N(0, 0.25) shifted to
    sum to zero, surface points uniform on the exposed union of (radius + 1.2 A) spheres, and
    the molecule's own Coulomb ESP at each surface point (COULOMB * sum q / d)."""
    heavy = random_walks(rng, [n_heavy])[0].astype(np.float64)
    nh = rng.integers(1, 3, size=n_heavy)
    hyd = np.concatenate([heavy[i] + H_BOND * _unit(rng, (nh[i],)) for i in range(n_heavy)])
    atoms = np.concatenate([heavy, hyd])
    radii = np.concatenate([np.full(n_heavy, VDW), np.full(len(hyd), VDW_H)])
    q = rng.normal(0.0, 0.25, size=len(atoms))
    q -= q.mean()
    rs = radii + SHAEP_PROBE
    got, need = [], n_surf
    while need > 0:
        m = max(2 * need, 64)
        which = rng.integers(0, len(atoms), size=m)
        pts = atoms[which] + rs[which, None] * _unit(rng, (m,))
        d = np.linalg.norm(pts[:, None, :] - atoms[None, :, :], axis=-1)
        d[np.arange(m), which] = np.inf
        keep = pts[(d >= rs[None, :] - 1e-9).all(axis=1)]
        got.append(keep[:need])
        need -= len(got[-1])
    surf = np.concatenate(got)
    esp = COULOMB * (q[None, :] / np.linalg.norm(surf[:, None, :] - atoms[None, :, :], axis=-1)).sum(1)
    f = lambda a: np.asarray(a, np.float32)
    return dict(atom_pos=f(heavy), centers_w_H=f(atoms), partial_charges=f(q), radii=f(radii),
                surf_pos=f(surf), surf_esp=f(esp))


def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (normalised 4-D Gaussian), (n, 4) float32."""
    q = rng.normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)


# ---------------------------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------------------------
def vol_pairs(n_pairs: int, lo: int = 20, hi: int = 40, seed: int = 0):
    """Atom-cloud pairs (ref list, fit list) with U{lo..hi} atoms each."""
    rng = np.random.default_rng(seed)
    ref = random_walks(rng, rng.integers(lo, hi + 1, size=n_pairs))
    fit = random_walks(rng, rng.integers(lo, hi + 1, size=n_pairs))
    return ref, fit


def surf_esp_pairs(n_pairs: int, atoms=(15, 50), points=(200, 400), seed: int = 1):
    """Surface + ESP pairs: (ref_pts, ref_esp, fit_pts, fit_esp) lists."""
    rng = np.random.default_rng(seed)
    mols = random_walks(rng, rng.integers(atoms[0], atoms[1] + 1, size=2 * n_pairs))
    npts = rng.integers(points[0], points[1] + 1, size=2 * n_pairs)
    pts = [surface_points(rng, m, int(k)) for m, k in zip(mols, npts)]
    esp = [esp_values(rng, len(p)) for p in pts]
    return pts[:n_pairs], esp[:n_pairs], pts[n_pairs:], esp[n_pairs:]


def pharm_shape_pairs(n_pairs: int, atoms=(5, 60), pharms=(2, 20), seed: int = 2):
    """Uneven atom + pharmacophore pairs: dict of ref/fit atoms, types, anchors, vectors."""
    rng = np.random.default_rng(seed)
    mols = random_walks(rng, rng.integers(atoms[0], atoms[1] + 1, size=2 * n_pairs))
    nph = rng.integers(pharms[0], pharms[1] + 1, size=2 * n_pairs)
    ph = [pharmacophores(rng, m, int(k)) for m, k in zip(mols, nph)]
    out = {}
    for side, sl in (("ref", slice(0, n_pairs)), ("fit", slice(n_pairs, 2 * n_pairs))):
        out[f"{side}_atoms"] = mols[sl]
        out[f"{side}_types"] = [p[0] for p in ph[sl]]
        out[f"{side}_anchors"] = [p[1] for p in ph[sl]]
        out[f"{side}_vectors"] = [p[2] for p in ph[sl]]
    return out


def conformer_library(n: int, atoms=(10, 50), base: int = 20000, seed: int = 5):
    """n atom clouds: ``base`` random-walk molecules, each posed under n/base uniform random
    rotations (the screening library; every entry is a distinct rigid conformer)."""
    rng = np.random.default_rng(seed)
    base = min(base, n)
    mols = random_walks(rng, rng.integers(atoms[0], atoms[1] + 1, size=base))
    q = random_quaternions(rng, n).astype(np.float64)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).reshape(n, 3, 3)
    which = np.arange(n) % base
    counts = np.array([len(mols[k]) for k in which], np.int64)
    flat = np.concatenate([mols[k] for k in which]).astype(np.float64)
    owner = np.repeat(np.arange(n), counts)
    flat = np.einsum("pij,pj->pi", R[owner], flat).astype(np.float32)
    return flat, counts


def screening_molecules(n: int, atoms=(10, 50), n_surf: int = 200, pharms=(1, 16), seed: int = 7):
    """n molecules of the screening workload: U{atoms} random-walk heavy atoms, n_surf surface
    points with ESP values, U{pharms} pharmacophores.  Returns a list of dicts with the
    container's attribute names (atom_pos, surf_pos, surf_esp, pharm_types, pharm_ancs,
    pharm_vecs)."""
    rng = np.random.default_rng(seed)
    mols = random_walks(rng, rng.integers(atoms[0], atoms[1] + 1, size=n))
    nph = rng.integers(pharms[0], pharms[1] + 1, size=n)
    out = []
    for m, k in zip(mols, nph):
        t, a, v = pharmacophores(rng, m, int(k))
        out.append(dict(atom_pos=m, surf_pos=surface_points(rng, m, n_surf), surf_esp=esp_values(rng, n_surf),
                        pharm_types=t.astype(np.int32), pharm_ancs=a, pharm_vecs=v))
    return out


def screening_library(n: int, device, atoms=(10, 50), n_surf: int = 200, pharms=(1, 16),
                      base: int = 8000, seed: int = 5, chunk: int = 65536):
    """The screening library: ``base`` molecules of :func:`screening_molecules`, each posed under
    n/base uniform random rotations (every entry a distinct rigid conformer).  Posing runs on
    the host in chunks of ``chunk`` entries (numpy); the packed rows are then copied to
    ``device``.  Returns (shape CloudSet of the heavy atoms, PharmSet, surface CloudSet with ESP
    in w)."""
    import torch
    from .packing import CloudSet, PharmSet
    mols = screening_molecules(min(base, n), atoms, n_surf, pharms, seed)
    base = len(mols)
    rng = np.random.default_rng(seed + 1)
    q = random_quaternions(rng, n).astype(np.float64)
    which = np.arange(n) % base

    def rotations(lo, hi):
        w, x, y, z = q[lo:hi].T
        return np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                         2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                         2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
                        1).reshape(hi - lo, 3, 3).astype(np.float32)

    def posed(key, w_key=None, rotate_only=False):
        """Packed (rows, 4) float32 of every entry's rows of ``key`` (entry i from base molecule
        i % base), rotated; w = the rows of ``w_key`` when given; and the per-entry counts."""
        bc = np.array([len(m[key]) for m in mols], np.int64)
        bstart = np.cumsum(bc) - bc
        brows = np.concatenate([np.asarray(m[key], np.float32).reshape(-1, 3) for m in mols])
        bw = np.concatenate([np.asarray(m[w_key], np.float32).reshape(-1) for m in mols]) if w_key else None
        counts = bc[which]
        first = np.cumsum(counts) - counts
        out = np.zeros((max(int(counts.sum()), 1), 4), np.float32)
        for lo in range(0, n, chunk):
            hi = min(n, lo + chunk)
            cnt = counts[lo:hi]
            owner = np.repeat(np.arange(hi - lo), cnt)
            local = np.arange(int(cnt.sum())) - (first[lo:hi] - first[lo])[owner]
            src = bstart[which[lo:hi]][owner] + local
            dst = slice(int(first[lo]), int(first[lo]) + int(cnt.sum()))
            out[dst, :3] = np.einsum("rij,rj->ri", rotations(lo, hi)[owner], brows[src])
            if bw is not None:
                out[dst, 3] = bw[src]
        return out, counts

    def rng_table(counts):
        return torch.from_numpy(np.stack([np.cumsum(counts) - counts, counts], 1).astype(np.int32)).to(device)

    dev = lambda a: torch.from_numpy(a).to(device)
    pts, ac = posed("atom_pos")
    shape = CloudSet(dev(pts), rng_table(ac), ac)
    del pts
    srf, sc = posed("surf_pos", "surf_esp")
    surface = CloudSet(dev(srf), rng_table(sc), sc)
    del srf
    anc, pc = posed("pharm_ancs")
    vec, _ = posed("pharm_vecs")
    tb = np.concatenate([m["pharm_types"] for m in mols]).astype(np.int32)
    bstart = np.cumsum([len(m["pharm_types"]) for m in mols]) - np.array([len(m["pharm_types"]) for m in mols])
    first = np.cumsum(pc) - pc
    owner = np.repeat(np.arange(n), pc)
    typ = tb[bstart[which][owner] + (np.arange(int(pc.sum())) - first[owner])]
    pharm = PharmSet(dev(anc), dev(vec), dev(np.ascontiguousarray(typ, np.int32)), rng_table(pc), pc)
    return shape, pharm, surface
