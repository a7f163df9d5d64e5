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


class ProfileStore:
    """Sharded on-disk store of conformer arrays; write once (create, add, close), then open."""

    def __init__(self, path: str, manifest: dict, mode: str):
        self.path, self.manifest, self._mode = path, manifest, mode
        self._buf: List[MoleculeProfile] = []
        self._n = int(manifest.get("n_total", 0))
        self._shard_id = len(manifest.get("shards", []))

    # -- writing ----------------------------------------------------------------------------
    @classmethod
    def create(cls, path: str, *, num_surf_points: int, modes: Sequence[str],
               dtype: str = "float16", shard_size: int = 100_000, pre_centered: bool = True,
               overwrite: bool = False) -> "ProfileStore":
        if dtype not in ("float16", "float32"):
            raise ValueError("dtype must be 'float16' or 'float32'")
        if shard_size < 1:
            raise ValueError("shard_size must be >= 1")
        schema = schema_from_modes(modes)
        os.makedirs(path, exist_ok=True)
        mpath = os.path.join(path, MANIFEST)
        if os.path.exists(mpath) and not overwrite:
            raise FileExistsError(f"{mpath} already exists; pass overwrite=True to replace it")
        if overwrite:      # only this store's own files
            for f in os.listdir(path):
                if f == MANIFEST or re.fullmatch(r"shard_\d{5}\.npz", f):
                    os.remove(os.path.join(path, f))
        manifest = dict(version=FORMAT_VERSION, num_surf_points=int(num_surf_points),
                        modes=list(modes), schema=schema, dtype=dtype, shard_size=int(shard_size),
                        pre_centered=bool(pre_centered), n_total=0, shards=[])
        store = cls(path, manifest, "w")
        store._write_manifest()
        return store

    def add(self, molecule, id=None) -> None:
        """Buffer one conformer (a MoleculeProfile or a container Molecule); a shard is written
        every ``shard_size`` conformers.  ``id`` defaults to the position in the store."""
        if self._mode != "w":
            raise RuntimeError("store is open for reading")
        prof = molecule if isinstance(molecule, MoleculeProfile) else MoleculeProfile.from_molecule(molecule)
        if id is None:
            id = prof.id if prof.id is not None else self._n + len(self._buf)
        self._buf.append(self._checked(prof, id))
        if len(self._buf) >= self.manifest["shard_size"]:
            self._flush()

    def add_profile(self, profile: "MoleculeProfile", id=None) -> None:
        """Add a MoleculeProfile as it is (screen.py:420-424); ``id`` defaults to its own."""
        if not isinstance(profile, MoleculeProfile):
            raise TypeError("add_profile expects a MoleculeProfile")
        self.add(profile, id=id if id is not None else profile.id)

    def _checked(self, p: MoleculeProfile, id) -> MoleculeProfile:
        sch = self.manifest["schema"]
        atom_pos = _f32(p.atom_pos).reshape(-1, 3)
        q = MoleculeProfile(atom_pos=atom_pos, id=id)
        if sch["surf"]:
            if p.surf_pos is None:
                raise ValueError("the store keeps surfaces but the conformer has none")
            q.surf_pos = _f32(p.surf_pos).reshape(-1, 3)
            if len(q.surf_pos) != self.manifest["num_surf_points"]:
                raise ValueError(f"surface has {len(q.surf_pos)} points, the store takes "
                                 f"{self.manifest['num_surf_points']}")
        if sch["surf_esp"]:
            if p.surf_esp is None:
                raise ValueError("the store keeps surface ESP but the conformer has none")
            q.surf_esp = _f32(p.surf_esp).reshape(-1)
        if sch["charges"]:
            if p.partial_charges is None:
                raise ValueError("the store keeps partial charges but the conformer has none")
            ch = _f32(p.partial_charges).reshape(-1)
            if sch["with_H"]:
                if p.nonH_atoms_idx is None or p.radii is None or p.centers_w_H is None:
                    raise ValueError("a with-H store needs nonH_atoms_idx, radii and centers_w_H")
                q.partial_charges = ch
                q.nonH_atoms_idx = np.asarray(p.nonH_atoms_idx, np.int64)
                q.radii = _f32(p.radii).reshape(-1)
                q.centers_w_H = _f32(p.centers_w_H).reshape(-1, 3)
            else:
                idx = np.asarray(p.nonH_atoms_idx, np.int64) if p.nonH_atoms_idx is not None \
                    else np.arange(len(ch))
                q.partial_charges = ch[idx]
            if p.atom_pos_noH is not None:
                q.atom_pos_noH = _f32(p.atom_pos_noH).reshape(-1, 3)
        if sch["pharm"]:
            if p.pharm_types is None:
                raise ValueError("the store keeps pharmacophores but the conformer has none")
            q.pharm_types = np.asarray(p.pharm_types, np.int32).reshape(-1)
            q.pharm_ancs = _f32(p.pharm_ancs).reshape(-1, 3)
            q.pharm_vecs = _f32(p.pharm_vecs).reshape(-1, 3)
        if self.manifest["pre_centered"]:
            mu = atom_pos.mean(0)
            q.atom_pos = q.atom_pos - mu
            for k in ("surf_pos", "pharm_ancs", "centers_w_H", "atom_pos_noH"):
                v = getattr(q, k)
                if v is not None:
                    setattr(q, k, v - mu)
        return q

    def close(self) -> None:
        if self._mode == "w":
            self._flush()
            self._write_manifest()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _write_manifest(self) -> None:
        tmp = os.path.join(self.path, MANIFEST + ".tmp")
        with open(tmp, "w") as fh:
            json.dump(self.manifest, fh, indent=2)
        os.replace(tmp, os.path.join(self.path, MANIFEST))

    def _flush(self) -> None:
        if not self._buf:
            return
        sch = self.manifest["schema"]
        dt = np.float16 if self.manifest["dtype"] == "float16" else np.float32
        recs = self._buf
        out = {"ids": np.array([r.id for r in recs])}
        out["atom_off"] = _offsets([len(r.atom_pos) for r in recs])
        out["atom_pos"] = np.concatenate([r.atom_pos for r in recs]).astype(dt)
        if sch["surf"]:
            out["surf_pos"] = np.stack([r.surf_pos for r in recs]).astype(dt)
        if sch["surf_esp"]:
            out["surf_esp"] = np.stack([r.surf_esp for r in recs]).astype(dt)
        if sch["charges"]:
            out["charges"] = np.concatenate([r.partial_charges for r in recs]).astype(dt)
            if sch["with_H"]:
                out["all_off"] = _offsets([len(r.partial_charges) for r in recs])
                out["nonH"] = np.concatenate([r.nonH_atoms_idx for r in recs]).astype(np.int64)
                out["radii"] = np.concatenate([r.radii for r in recs]).astype(dt)
                out["cwh"] = np.concatenate([r.centers_w_H for r in recs]).astype(dt)
                heavy = [len(r.nonH_atoms_idx) for r in recs]
            else:
                heavy = [len(r.partial_charges) for r in recs]
            if any(r.atom_pos_noH is not None for r in recs):
                out["heavy_off"] = _offsets(heavy)
                out["xyz_noH"] = np.concatenate([r.atom_pos_noH if r.atom_pos_noH is not None
                                                 else r.atom_pos for r in recs]).astype(dt)
        if sch["pharm"]:
            out["pharm_off"] = _offsets([len(r.pharm_types) for r in recs])
            out["pharm_types"] = np.concatenate([r.pharm_types for r in recs]).astype(np.int32)
            out["pharm_ancs"] = np.concatenate([r.pharm_ancs for r in recs]).astype(dt)
            out["pharm_vecs"] = np.concatenate([r.pharm_vecs for r in recs]).astype(dt)
        name = f"shard_{self._shard_id:05d}.npz"
        np.savez(os.path.join(self.path, name), **out)
        self.manifest["shards"].append(dict(name=name, n=len(recs), start=self._n))
        self._n += len(recs)
        self.manifest["n_total"] = self._n
        self._buf = []
        self._shard_id += 1
        self._write_manifest()

    # -- reading ----------------------------------------------------------------------------
    @classmethod
    def open(cls, path: str) -> "ProfileStore":
        with open(os.path.join(path, MANIFEST)) as fh:
            manifest = json.load(fh)
        ver = manifest.get("version")
        if ver != FORMAT_VERSION:
            raise ValueError(f"store at {path!r} has on-disk format version {ver!r}; this reader "
                             f"takes version {FORMAT_VERSION}. Rebuild the store.")
        return cls(path, manifest, "r")

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
        """(shard metadata, {array name: ndarray}) of shard ``idx``; no pickled objects."""
        sh = self.manifest["shards"][idx]
        try:
            with np.load(os.path.join(self.path, sh["name"]), allow_pickle=False) as data:
                return sh, {k: data[k] for k in data.files}
        except ValueError as exc:
            raise ValueError(f"{sh['name']}: arrays that need unpickling (object ids?) are not "
                             f"read: {exc}") from exc

    def iter_profiles(self) -> Iterator[List[MoleculeProfile]]:
        """One shard at a time as MoleculeProfile lists (float32 in memory)."""
        for i in range(self.num_shards):
            yield self.read_profiles(i)

    iter_shards = iter_profiles            # the reference's name (screen.py:556-560)

    def read_profiles(self, idx: int) -> List[MoleculeProfile]:
        """Shard ``idx`` as a MoleculeProfile list (random access; screen.py:562-565)."""
        sch = self.schema
        sh, a = self.read_shard(idx)
        ao = a["atom_off"]
        po = a.get("pharm_off")
        out = []
        for j in range(sh["n"]):
            p = MoleculeProfile(atom_pos=_f32(a["atom_pos"][ao[j]:ao[j + 1]]), id=a["ids"][j].item())
            if sch["surf"]:
                p.surf_pos = _f32(a["surf_pos"][j])
            if sch["surf_esp"]:
                p.surf_esp = _f32(a["surf_esp"][j])
            if sch["charges"]:
                if sch["with_H"]:
                    c0, c1 = a["all_off"][j], a["all_off"][j + 1]
                    p.partial_charges = _f32(a["charges"][c0:c1])
                    p.nonH_atoms_idx = a["nonH"][ao[j]:ao[j + 1]]
                    p.radii = _f32(a["radii"][c0:c1])
                    p.centers_w_H = _f32(a["cwh"][c0:c1])
                else:
                    ho = a["heavy_off"] if "heavy_off" in a else ao
                    p.partial_charges = _f32(a["charges"][ho[j]:ho[j + 1]])
            if sch["pharm"]:
                p.pharm_types = a["pharm_types"][po[j]:po[j + 1]]
                p.pharm_ancs = _f32(a["pharm_ancs"][po[j]:po[j + 1]])
                p.pharm_vecs = _f32(a["pharm_vecs"][po[j]:po[j + 1]])
            out.append(p)
        return out


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
    return out
