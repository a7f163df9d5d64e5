"""On-disk conformer store (the reference's ``ProfileStore`` format, version 1) and a loader that
streams its shards straight into an HBM-resident native library.

Format (restated from shepherd_score/screen.py:317-605; it is an interchange format, so stores
written by either implementation read in the other)::

    <path>/manifest.json      version, num_surf_points, modes, schema, dtype, shard_size,
                              pre_centered, n_total, shards [{name, n, start}]
    <path>/shard_NNNNN.npz    ids, atom_off (CSR int64) + atom_pos, surf_pos (K, S, 3),
                              surf_esp (K, S), charges [+ all_off, nonH, radii, cwh for with-H
                              stores; heavy_off + xyz_noH when heavy centres differ from
                              atom_pos], pharm_off + pharm_types (int32) / pharm_ancs / pharm_vecs

Coordinates and charges are float16 (default) or float32 on disk; every loader upcasts to float32.
``pre_centered`` stores shift each profile by its ``atom_pos`` mean at write time.

:func:`load_library` reads a store shard by shard (``numpy.load`` without pickle), builds the
float4 rows the kernels consume (ESP / charge in ``w``) with vectorised numpy, and appends them to
a native ``FssLibrary`` (``fss_lib_upload``) -- no per-molecule Python loop and no torch copy.
"""
from __future__ import annotations

import ctypes
import json
import os
import re
from dataclasses import dataclass
from typing import Iterator, List, Optional, Sequence

import numpy as np

from . import _lib
from .modes import canonical

FORMAT_VERSION = 1
MANIFEST = "manifest.json"
STORE_MODES = ("vol", "vol_esp", "surf", "surf_esp", "pharm", "vol_color", "vol_and_surf_esp")


def schema_from_modes(modes: Sequence[str]) -> dict:
    """Which arrays a store keeps for the given modes (screen.py:203-217)."""
    ms = {canonical(m) for m in modes}
    unknown = ms - set(STORE_MODES)
    if unknown:
        raise ValueError(f"unknown modes {sorted(unknown)}; valid: {list(STORE_MODES)}")
    full = "vol_and_surf_esp" in ms
    return dict(surf=bool({"surf", "surf_esp", "vol_and_surf_esp"} & ms),
                surf_esp=bool({"surf_esp", "vol_and_surf_esp"} & ms),
                charges=bool({"vol_esp", "vol_and_surf_esp"} & ms),
                with_H=full, radii=full, centers_w_H=full,
                pharm=bool({"pharm", "vol_color"} & ms))


def store_supports(schema: dict, mode: str) -> bool:
    """Whether a store with ``schema`` holds what ``mode`` reads (screen.py:220-238)."""
    m = canonical(mode)
    if m == "vol":
        return True
    if m == "vol_esp":
        return bool(schema["charges"])
    if m == "surf":
        return bool(schema["surf"])
    if m == "surf_esp":
        return bool(schema["surf"] and schema["surf_esp"])
    if m in ("pharm", "vol_color"):
        return bool(schema["pharm"])
    if m == "vol_and_surf_esp":
        return all(schema[k] for k in ("surf", "surf_esp", "centers_w_H", "radii", "charges", "with_H"))
    return False


@dataclass
class MoleculeProfile:
    """The arrays of one conformer that a store keeps (no RDKit object)."""
    atom_pos: np.ndarray
    surf_pos: Optional[np.ndarray] = None
    surf_esp: Optional[np.ndarray] = None
    partial_charges: Optional[np.ndarray] = None
    radii: Optional[np.ndarray] = None
    nonH_atoms_idx: Optional[np.ndarray] = None
    pharm_types: Optional[np.ndarray] = None
    pharm_ancs: Optional[np.ndarray] = None
    pharm_vecs: Optional[np.ndarray] = None
    centers_w_H: Optional[np.ndarray] = None
    atom_pos_noH: Optional[np.ndarray] = None
    id: object = None

    def center_to(self, xyz_means) -> None:
        """Move every position array by -xyz_means; vectors, charges and radii are unchanged
        (screen.py:178-190)."""
        mu = np.asarray(xyz_means, dtype=np.float32)
        for k in ("atom_pos", "atom_pos_noH", "surf_pos", "pharm_ancs", "centers_w_H"):
            v = getattr(self, k)
            if v is not None:
                setattr(self, k, np.asarray(v, np.float32) - mu)

    @classmethod
    def from_molecule(cls, m, *, modes=None, id=None) -> "MoleculeProfile":
        """From anything with the container's Molecule attributes (missing ones stay None).
        ``modes``: keep only the arrays those modes need (screen.py:193-197)."""
        g = lambda k: getattr(m, k, None)
        p = cls(atom_pos=np.asarray(m.atom_pos, np.float32), surf_pos=g("surf_pos"),
                surf_esp=g("surf_esp"), partial_charges=g("atom_charges") if g("partial_charges") is None
                else g("partial_charges"), radii=g("radii"), nonH_atoms_idx=g("nonH_atoms_idx"),
                pharm_types=g("pharm_types"), pharm_ancs=g("pharm_ancs"), pharm_vecs=g("pharm_vecs"),
                centers_w_H=g("centers_w_H"), atom_pos_noH=g("atom_pos_noH"),
                id=id if id is not None else g("id"))
        if modes is not None:
            sch = schema_from_modes(modes)
            drop = []
            if not sch["surf"]:
                drop += ["surf_pos"]
            if not sch["surf_esp"]:
                drop += ["surf_esp"]
            if not sch["charges"]:
                drop += ["partial_charges", "nonH_atoms_idx"]
            if not sch["with_H"]:
                drop += ["radii", "centers_w_H"]
            if not sch["pharm"]:
                drop += ["pharm_types", "pharm_ancs", "pharm_vecs"]
            for k in drop:
                setattr(p, k, None)
        return p


def _f32(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float32)


def _offsets(lengths) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(np.asarray(lengths, np.int64))]).astype(np.int64)


# Columns of one conformer as a version-1 shard stores them.  A shard concatenates each ragged
# column behind a CSR offsets array and stacks the fixed-size surface columns:
#   atom_off   atom_pos, and nonH (the heavy-atom indices) in with-H stores
#   all_off    charges, radii, cwh (every atom)          with-H stores
#   heavy_off  xyz_noH                                   when a conformer keeps separate heavy centres
#   pharm_off  pharm_types (int32), pharm_ancs, pharm_vecs
#   (stacked)  surf_pos (n, S, 3), surf_esp (n, S)
# Without hydrogens, `charges` holds the heavy-atom charges and follows heavy_off when it exists,
# atom_off otherwise.  Coordinates and charges take the store's dtype; offsets and nonH are int64.
_POSITIONS = ("atom_pos", "surf_pos", "pharm_ancs", "cwh", "xyz_noH")


def _shard_file(k: int) -> str:
    return f"shard_{k:05d}.npz"


_SHARD_FILE = re.compile(r"shard_\d{5}\.npz")


def _encode_shard(ids: list, cols: List[dict], schema: dict, dtype) -> dict:
    """The npz arrays of one shard from its conformers' column dicts."""
    def cat(key, dt=dtype):
        return np.concatenate([c[key] for c in cols]).astype(dt)

    def off(key):
        return _offsets([len(c[key]) for c in cols])

    arrays = {"ids": np.asarray(ids), "atom_off": off("atom_pos"), "atom_pos": cat("atom_pos")}
    for flag, key in (("surf", "surf_pos"), ("surf_esp", "surf_esp")):
        if schema[flag]:
            arrays[key] = np.stack([c[key] for c in cols]).astype(dtype)
    if schema["charges"]:
        arrays["charges"] = cat("charges")
        if schema["with_H"]:
            arrays["all_off"] = off("charges")
            arrays["nonH"] = cat("nonH", np.int64)
            arrays["radii"], arrays["cwh"] = cat("radii"), cat("cwh")
        if any("xyz_noH" in c for c in cols):
            heavy = [c.get("xyz_noH", c["atom_pos"]) for c in cols]
            arrays["heavy_off"] = _offsets([len(c["nonH"] if schema["with_H"] else c["charges"])
                                            for c in cols])
            arrays["xyz_noH"] = np.concatenate(heavy).astype(dtype)
    if schema["pharm"]:
        arrays["pharm_off"] = off("pharm_types")
        arrays["pharm_types"] = cat("pharm_types", np.int32)
        arrays["pharm_ancs"], arrays["pharm_vecs"] = cat("pharm_ancs"), cat("pharm_vecs")
    return arrays


def _decode_shard(a: dict, n: int, schema: dict) -> List[MoleculeProfile]:
    """The MoleculeProfile list of one shard's arrays (float32 in memory)."""
    def spans(key):
        o = a[key]
        return [slice(int(o[j]), int(o[j + 1])) for j in range(n)]

    atoms = spans("atom_off")
    every = spans("all_off") if schema["charges"] and schema["with_H"] else None
    heavy = spans("heavy_off") if "heavy_off" in a else atoms
    feats = spans("pharm_off") if schema["pharm"] else None
    out = []
    for j in range(n):
        p = MoleculeProfile(atom_pos=_f32(a["atom_pos"][atoms[j]]), id=a["ids"][j].item())
        if schema["surf"]:
            p.surf_pos = _f32(a["surf_pos"][j])
        if schema["surf_esp"]:
            p.surf_esp = _f32(a["surf_esp"][j])
        if every is not None:
            e = every[j]
            p.partial_charges, p.radii, p.centers_w_H = (_f32(a[k][e]) for k in ("charges", "radii", "cwh"))
            p.nonH_atoms_idx = a["nonH"][atoms[j]]
        elif schema["charges"]:
            p.partial_charges = _f32(a["charges"][heavy[j]])
        if feats is not None:
            f = feats[j]
            p.pharm_types = a["pharm_types"][f]
            p.pharm_ancs, p.pharm_vecs = _f32(a["pharm_ancs"][f]), _f32(a["pharm_vecs"][f])
        out.append(p)
    return out


class ProfileStore:
    """A sharded on-disk conformer store in the version-1 interchange format.

    Written once -- ``create``, ``add`` / ``add_profile``, ``close`` (or a ``with`` block) -- and
    read through ``open``.  Conformers wait in column form until a shard's worth
    (``shard_size``) is buffered; each shard write is followed by an atomic manifest update, so
    an interrupted writer leaves a readable store of the shards already written.
    """

    def __init__(self, path: str, manifest: dict, writable: bool):
        self.path = path
        self.manifest = manifest
        self._writable = writable
        self._pending: List[dict] = []        # column dicts not yet in a shard
        self._pending_ids: list = []

    # -- writing ----------------------------------------------------------------------------
    @classmethod
    def create(cls, path: str, *, num_surf_points: int, modes: Sequence[str],
               dtype: str = "float16", shard_size: int = 100_000, pre_centered: bool = True,
               overwrite: bool = False) -> "ProfileStore":
        """A new empty store at ``path`` for ``modes``; ``overwrite`` replaces an existing
        store there (its manifest and shard files only; other files in the folder stay)."""
        if dtype not in ("float16", "float32"):
            raise ValueError(f"unsupported store dtype {dtype!r}: use 'float16' or 'float32'")
        if int(shard_size) < 1:
            raise ValueError(f"shard_size must be positive, got {shard_size}")
        schema = schema_from_modes(modes)
        os.makedirs(path, exist_ok=True)
        if os.path.exists(os.path.join(path, MANIFEST)):
            if not overwrite:
                raise FileExistsError(f"a store already exists at {path!r} (overwrite=True replaces it)")
            for f in os.listdir(path):
                if f == MANIFEST or _SHARD_FILE.fullmatch(f):
                    os.remove(os.path.join(path, f))
        store = cls(path, {"version": FORMAT_VERSION, "num_surf_points": int(num_surf_points),
                           "modes": list(modes), "schema": schema, "dtype": dtype,
                           "shard_size": int(shard_size), "pre_centered": bool(pre_centered),
                           "n_total": 0, "shards": []}, writable=True)
        store._save_manifest()
        return store

    def add(self, molecule, id=None) -> None:
        """Buffer one conformer (a MoleculeProfile, or anything with the container Molecule's
        attributes); ``id`` defaults to the profile's own, else the conformer's position."""
        if not self._writable:
            raise RuntimeError(f"the store at {self.path!r} was opened for reading; create() a new one to write")
        prof = molecule if isinstance(molecule, MoleculeProfile) else MoleculeProfile.from_molecule(molecule)
        if id is None:
            id = prof.id if prof.id is not None else self.manifest["n_total"] + len(self._pending)
        self._pending.append(self._columns(prof, id))
        self._pending_ids.append(id)
        if len(self._pending) == self.manifest["shard_size"]:
            self._write_shard()

    def add_profile(self, profile: "MoleculeProfile", id=None) -> None:
        """Add a MoleculeProfile unchanged (screen.py:420-424); ``id`` defaults to its own."""
        if not isinstance(profile, MoleculeProfile):
            raise TypeError(f"add_profile takes a MoleculeProfile, not {type(profile).__name__}")
        self.add(profile, id=id)

    def _columns(self, p: MoleculeProfile, ident) -> dict:
        """Validate a profile against the schema and return its shard columns (float32 here;
        the store's dtype is applied when the shard is written)."""
        schema, S = self.schema, self.num_surf_points

        def required(attr, what):
            v = getattr(p, attr)
            if v is None:
                raise ValueError(f"conformer {ident!r} has no {attr}, which this store keeps ({what})")
            return v

        xyz = _f32(p.atom_pos).reshape(-1, 3)
        col = {"atom_pos": xyz}
        if schema["surf"]:
            col["surf_pos"] = _f32(required("surf_pos", "surfaces")).reshape(-1, 3)
            if len(col["surf_pos"]) != S:
                raise ValueError(f"conformer {ident!r} has {len(col['surf_pos'])} surface points; "
                                 f"the store holds {S} points per surface")
        if schema["surf_esp"]:
            col["surf_esp"] = _f32(required("surf_esp", "surface ESP")).reshape(-1)
        if schema["charges"]:
            q = _f32(required("partial_charges", "partial charges")).reshape(-1)
            if schema["with_H"]:
                col["charges"] = q
                col["nonH"] = np.asarray(required("nonH_atoms_idx", "with-H rows"), np.int64).reshape(-1)
                col["radii"] = _f32(required("radii", "with-H rows")).reshape(-1)
                col["cwh"] = _f32(required("centers_w_H", "with-H rows")).reshape(-1, 3)
            else:
                pick = p.nonH_atoms_idx
                col["charges"] = q if pick is None else q[np.asarray(pick, np.int64)]
            if p.atom_pos_noH is not None:
                col["xyz_noH"] = _f32(p.atom_pos_noH).reshape(-1, 3)
        if schema["pharm"]:
            col["pharm_types"] = np.asarray(required("pharm_types", "pharmacophores"), np.int32).reshape(-1)
            col["pharm_ancs"] = _f32(p.pharm_ancs).reshape(-1, 3)
            col["pharm_vecs"] = _f32(p.pharm_vecs).reshape(-1, 3)
        if self.pre_centered:          # shift every position column by the heavy-atom mean
            mu = xyz.mean(0)
            for key in _POSITIONS:
                if key in col:
                    col[key] = col[key] - mu
        return col

    def _write_shard(self) -> None:
        if not self._pending:
            return
        shards = self.manifest["shards"]
        name = _shard_file(len(shards))
        dt = np.float16 if self.manifest["dtype"] == "float16" else np.float32
        np.savez(os.path.join(self.path, name),
                 **_encode_shard(self._pending_ids, self._pending, self.schema, dt))
        n = len(self._pending)
        shards.append({"name": name, "n": n, "start": self.manifest["n_total"]})
        self.manifest["n_total"] += n
        self._pending, self._pending_ids = [], []
        self._save_manifest()

    def _save_manifest(self) -> None:
        part = os.path.join(self.path, MANIFEST + ".tmp")
        with open(part, "w") as fh:
            json.dump(self.manifest, fh, indent=2)
        os.replace(part, os.path.join(self.path, MANIFEST))

    def close(self) -> None:
        """Write the buffered conformers as a last (short) shard; reading stores ignore it."""
        if self._writable:
            self._write_shard()
            self._save_manifest()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reading ----------------------------------------------------------------------------
    @classmethod
    def open(cls, path: str) -> "ProfileStore":
        with open(os.path.join(path, MANIFEST)) as fh:
            manifest = json.load(fh)
        found = manifest.get("version")
        if found != FORMAT_VERSION:
            raise ValueError(f"{path!r}: manifest format version {found!r}, this package reads version "
                             f"{FORMAT_VERSION}; write the store again with this version")
        return cls(path, manifest, writable=False)

    def __len__(self) -> int:
        return int(self.manifest["n_total"])

    @property
    def num_surf_points(self) -> int:
        return int(self.manifest["num_surf_points"])

    @property
    def modes(self) -> tuple:
        return tuple(self.manifest["modes"])

    @property
    def schema(self) -> dict:
        return self.manifest["schema"]

    @property
    def pre_centered(self) -> bool:
        return bool(self.manifest["pre_centered"])

    @property
    def num_shards(self) -> int:
        return len(self.manifest["shards"])

    def supports(self, mode: str) -> bool:
        return store_supports(self.schema, mode)

    def read_shard(self, idx: int):
        """(shard record, {array name: ndarray}) of shard ``idx``; object arrays are refused."""
        rec = self.manifest["shards"][idx]
        try:
            with np.load(os.path.join(self.path, rec["name"]), allow_pickle=False) as z:
                return rec, {k: z[k] for k in z.files}
        except ValueError as exc:
            raise ValueError(f"{rec['name']} holds arrays that only unpickling could read "
                             f"(object ids?); they are not loaded: {exc}") from exc

    def read_profiles(self, idx: int) -> List[MoleculeProfile]:
        """Shard ``idx`` as a MoleculeProfile list, random access (screen.py:562-565)."""
        rec, arrays = self.read_shard(idx)
        return _decode_shard(arrays, int(rec["n"]), self.schema)

    def iter_profiles(self) -> Iterator[List[MoleculeProfile]]:
        """The shards in order, each as a MoleculeProfile list (float32 in memory)."""
        return (self.read_profiles(i) for i in range(self.num_shards))

    iter_shards = iter_profiles            # the reference's name (screen.py:556-560)


def _rows(xyz: np.ndarray, w: Optional[np.ndarray] = None) -> np.ndarray:
    r = np.zeros((len(xyz), 4), np.float32)
    r[:, :3] = xyz
    if w is not None:
        r[:, 3] = w
    return r


def _shard_channels(a: dict, mode: str, schema: dict):
    """(shape rows, counts) and (anc rows, vec rows, types, counts) of one shard for ``mode``."""
    shape = pharm = None
    if mode in ("vol", "vol_color", "vol_and_surf_esp"):   # ShaEP: heavy atoms for alpha 0.81
        ao = a["atom_off"]
        shape = _rows(_f32(a["atom_pos"])), np.diff(ao).astype(np.int32)
    elif mode == "vol_esp":
        ao = a["atom_off"]
        ho = a["heavy_off"] if "heavy_off" in a else ao
        xyz = _f32(a["xyz_noH"]) if "xyz_noH" in a else _f32(a["atom_pos"])
        if "all_off" in a:   # with-H store: heavy charges = charges[all_off start + nonH]
            ch = a["charges"][a["nonH"] + np.repeat(a["all_off"][:-1], np.diff(ho))]
        else:
            ch = a["charges"]
        shape = _rows(xyz, _f32(ch)), np.diff(ho).astype(np.int32)
    elif mode in ("surf", "surf_esp"):
        sp = _f32(a["surf_pos"])
        K, S = sp.shape[:2]
        w = _f32(a["surf_esp"]).reshape(-1) if mode == "surf_esp" else None
        shape = _rows(sp.reshape(-1, 3), w), np.full(K, S, np.int32)
    if mode in ("pharm", "vol_color"):
        po = a["pharm_off"]
        pharm = (_rows(_f32(a["pharm_ancs"])), _rows(_f32(a["pharm_vecs"])),
                 np.ascontiguousarray(a["pharm_types"], np.int32), np.diff(po).astype(np.int32))
    return shape, pharm


def _shard_shaep(a: dict):
    """ShaEP channel of one with-H shard: (surface rows (x, y, z, ESP), surface counts, atom rows
    (x, y, z, partial charge) with hydrogens, radii, atom counts)"""
    sp = _f32(a["surf_pos"])
    K, S = sp.shape[:2]
    srf = _rows(sp.reshape(-1, 3), _f32(a["surf_esp"]).reshape(-1))
    atm = _rows(_f32(a["cwh"]), _f32(a["charges"]))
    return (srf, np.full(K, S, np.int32), atm, np.ascontiguousarray(_f32(a["radii"])),
            np.diff(a["all_off"]).astype(np.int32))


def _device_index(device) -> int:
    if isinstance(device, int):
        return device
    import torch
    d = torch.device(device)
    return d.index if d.index is not None else torch.cuda.current_device()


def _center_shard(a: dict) -> dict:
    """A shard's arrays with every conformer moved so that its ``atom_pos`` mean is the origin
    (what a pre-centred store does at write time)."""
    ao = a["atom_off"]
    cnt = np.diff(ao)
    mu = np.add.reduceat(_f32(a["atom_pos"]).astype(np.float64), ao[:-1], axis=0) / np.maximum(cnt, 1)[:, None]
    mu[cnt == 0] = 0.0
    out = dict(a)
    per = lambda off: np.repeat(mu, np.diff(off), axis=0)
    out["atom_pos"] = (_f32(a["atom_pos"]) - per(ao)).astype(np.float32)
    if "xyz_noH" in a:
        out["xyz_noH"] = (_f32(a["xyz_noH"]) - per(a["heavy_off"])).astype(np.float32)
    if "surf_pos" in a:
        out["surf_pos"] = (_f32(a["surf_pos"]) - mu[:, None, :]).astype(np.float32)
    if "pharm_ancs" in a:
        out["pharm_ancs"] = (_f32(a["pharm_ancs"]) - per(a["pharm_off"])).astype(np.float32)
    if "cwh" in a:
        out["cwh"] = (_f32(a["cwh"]) - per(a["all_off"])).astype(np.float32)
    return out


def load_library(store, mode: str, device=0, shards: Optional[Sequence[int]] = None,
                 center: bool = False):
    """Stream ``store`` (a ProfileStore or its path) into a native HBM library for ``mode``;
    returns a ConformerLibrary whose rows live only in the native library.  ``center``: a store
    that is not pre-centred has each conformer centred on its ``atom_pos`` mean on the way in."""
    from .screen import ConformerLibrary
    if isinstance(store, (str, os.PathLike)):
        store = ProfileStore.open(str(store))
    m = canonical(mode)
    if not store.supports(m):
        raise ValueError(f"store does not support mode {m!r} (modes {store.modes})")
    idx = list(range(store.num_shards)) if shards is None else list(shards)
    want_shape = m != "pharm"
    want_pharm = m in ("pharm", "vol_color")
    want_se = m == "vol_and_surf_esp"
    # sizing pass: conformers and rows per channel from the offsets alone
    n = rows = prows = srows = arows = 0
    for i in idx:
        sh = store.manifest["shards"][i]
        path = os.path.join(store.path, sh["name"])
        with np.load(path, allow_pickle=False) as d:
            n += int(sh["n"])
            if want_shape:
                if m in ("surf", "surf_esp"):
                    rows += int(sh["n"]) * store.num_surf_points
                elif m == "vol_esp" and "heavy_off" in d.files:
                    rows += int(d["heavy_off"][-1])
                else:
                    rows += int(d["atom_off"][-1])
            if want_pharm:
                prows += int(d["pharm_off"][-1])
            if want_se:
                srows += int(sh["n"]) * store.num_surf_points
                arows += int(d["all_off"][-1])
    lib = _lib.load()
    spec = _lib.FssLibSpec()
    spec.device = _device_index(device)
    spec.channels = ((_lib.LIB_SHAPE if want_shape else 0) | (_lib.LIB_PHARM if want_pharm else 0) |
                     (_lib.LIB_SHAEP if want_se else 0))
    spec.max_conformers = max(n, 1)
    spec.max_points = max(rows, 1) if want_shape else 0
    spec.max_pharm = max(prows, 1) if want_pharm else 0
    spec.max_surface = max(srows, 1) if want_se else 0
    spec.max_atoms = max(arows, 1) if want_se else 0
    h = ctypes.c_void_p()
    _lib.call("fss_lib_create", _lib.byref(spec), ctypes.byref(h))
    ids = []
    nbytes = 0
    first = ctypes.c_int64()
    try:
        recenter = center and not store.pre_centered
        for i in idx:
            sh, a = store.read_shard(i)
            if recenter:
                a = _center_shard(a)
            shape, pharm = _shard_channels(a, m, store.schema)
            r = _lib.FssLibRows()
            if shape is not None:
                r.pts, r.counts = shape[0].ctypes.data, shape[1].ctypes.data
                nbytes += shape[0].nbytes
            if pharm is not None:
                r.anc, r.vec, r.typ, r.pcounts = (pharm[0].ctypes.data, pharm[1].ctypes.data,
                                                  pharm[2].ctypes.data, pharm[3].ctypes.data)
                nbytes += pharm[0].nbytes + pharm[1].nbytes + pharm[2].nbytes
            if want_se:
                se = _shard_shaep(a)
                r.srf, r.scounts, r.atm, r.rad, r.acounts = (x.ctypes.data for x in se)
                nbytes += se[0].nbytes + se[2].nbytes + se[3].nbytes
            _lib.call("fss_lib_append", h, int(sh["n"]), _lib.byref(r), ctypes.byref(first))
            ids.append(np.asarray(a["ids"]))
    except Exception:
        lib.fss_lib_destroy(h)
        raise
    all_ids = np.concatenate(ids) if ids else np.zeros(0, np.int64)
    out = ConformerLibrary.from_native(h, n, all_ids, nbytes)
    out.num_surf_points = store.num_surf_points if store.schema.get("surf") else None
    out.pre_centered = store.pre_centered or recenter
    out._native_surface = want_se
    return out
