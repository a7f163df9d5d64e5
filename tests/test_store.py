"""The on-disk conformer store (ProfileStore format version 1) and the loader's row building:
round trips, schema rules, version check, the exact on-disk key set per schema (the interchange
contract of shepherd_score/screen.py:317-605), and the vectorised float4 rows against a
per-molecule reconstruction.  No GPU."""
import json
import os

import numpy as np
import pytest

from fast_shepherd_score_amd.store import (FORMAT_VERSION, MoleculeProfile, ProfileStore,
                                           _shard_channels, schema_from_modes, store_supports)

ALL_MODES = ("vol", "vol_esp", "surf", "surf_esp", "pharm", "vol_color", "vol_and_surf_esp")


def _profile(seed, with_h=False, S=32, heavy=6, total=9, P=5, noH=False):
    rng = np.random.default_rng(seed)
    p = MoleculeProfile(atom_pos=rng.normal(size=(heavy, 3)).astype(np.float32) + seed,
                        surf_pos=rng.normal(size=(S, 3)).astype(np.float32),
                        surf_esp=rng.normal(size=S).astype(np.float32) * 0.1,
                        pharm_types=rng.integers(0, 8, size=P).astype(np.int32),
                        pharm_ancs=rng.normal(size=(P, 3)).astype(np.float32),
                        pharm_vecs=rng.normal(size=(P, 3)).astype(np.float32), id=100 + seed)
    if with_h:
        p.partial_charges = rng.normal(size=total).astype(np.float32)
        p.nonH_atoms_idx = np.sort(rng.choice(total, size=heavy, replace=False)).astype(np.int64)
        p.radii = (1 + rng.random(total)).astype(np.float32)
        p.centers_w_H = rng.normal(size=(total, 3)).astype(np.float32)
    else:
        p.partial_charges = rng.normal(size=heavy).astype(np.float32)
    if noH:
        p.atom_pos_noH = p.atom_pos + 0.25
    return p


def test_roundtrip_every_mode_float32(tmp_path):
    path = os.path.join(tmp_path, "lib.fss")
    profs = [_profile(i, with_h=True) for i in range(7)]
    with ProfileStore.create(path, num_surf_points=32, modes=ALL_MODES, dtype="float32",
                             shard_size=3, pre_centered=False) as st:
        for p in profs:
            st.add(p)
    st = ProfileStore.open(path)
    assert len(st) == 7 and st.num_shards == 3 and st.num_surf_points == 32
    assert all(st.supports(m) for m in ALL_MODES + ("esp", "esp_combo"))
    got = [p for shard in st.iter_profiles() for p in shard]
    for a, b in zip(profs, got):
        assert b.id == a.id
        for k in ("atom_pos", "surf_pos", "surf_esp", "partial_charges", "nonH_atoms_idx", "radii",
                  "pharm_types", "pharm_ancs", "pharm_vecs", "centers_w_H"):
            np.testing.assert_array_equal(getattr(b, k), getattr(a, k))


def test_heavy_only_charges_and_identity_selection(tmp_path):
    path = os.path.join(tmp_path, "lib.fss")
    profs = [_profile(i) for i in range(4)]
    with ProfileStore.create(path, num_surf_points=32, modes=("vol", "vol_esp"), dtype="float32",
                             pre_centered=False) as st:
        for p in profs:
            st.add(p)
    st = ProfileStore.open(path)
    assert not st.schema["with_H"] and not st.supports("surf_esp") and st.supports("vol_esp")
    for a, b in zip(profs, next(st.iter_profiles())):
        np.testing.assert_array_equal(b.partial_charges, a.partial_charges)


def test_pre_centered_shifts_every_array_by_the_atom_mean(tmp_path):
    path = os.path.join(tmp_path, "lib.fss")
    p = _profile(3)
    with ProfileStore.create(path, num_surf_points=32, modes=("vol", "surf", "pharm"),
                             dtype="float32", pre_centered=True) as st:
        st.add(p)
    g = next(ProfileStore.open(path).iter_profiles())[0]
    mu = p.atom_pos.mean(0)
    np.testing.assert_allclose(g.atom_pos.mean(0), 0.0, atol=1e-5)
    np.testing.assert_allclose(g.surf_pos, p.surf_pos - mu, atol=1e-6)
    np.testing.assert_allclose(g.pharm_ancs, p.pharm_ancs - mu, atol=1e-6)
    np.testing.assert_array_equal(g.pharm_vecs, p.pharm_vecs)      # directions are not shifted


def test_float16_is_lossy_but_close_and_upcast(tmp_path):
    path = os.path.join(tmp_path, "lib.fss")
    p = _profile(1)
    with ProfileStore.create(path, num_surf_points=32, modes=("vol", "surf"), dtype="float16",
                             pre_centered=False) as st:
        st.add(p)
    g = next(ProfileStore.open(path).iter_profiles())[0]
    assert g.atom_pos.dtype == np.float32
    np.testing.assert_allclose(g.atom_pos, p.atom_pos, atol=2e-2)
    np.testing.assert_array_equal(g.atom_pos, p.atom_pos.astype(np.float16).astype(np.float32))


def test_version_check_and_overwrite(tmp_path):
    path = os.path.join(tmp_path, "lib.fss")
    with ProfileStore.create(path, num_surf_points=32, modes=("vol",)) as st:
        st.add(_profile(0))
    with pytest.raises(FileExistsError):
        ProfileStore.create(path, num_surf_points=32, modes=("vol",))
    open(os.path.join(path, "keep.npz"), "wb").close()         # not a store file
    st = ProfileStore.create(path, num_surf_points=32, modes=("vol",), overwrite=True)
    st.close()
    assert os.path.exists(os.path.join(path, "keep.npz"))
    assert len(ProfileStore.open(path)) == 0
    m = json.load(open(os.path.join(path, "manifest.json")))
    m["version"] = FORMAT_VERSION + 1
    json.dump(m, open(os.path.join(path, "manifest.json"), "w"))
    with pytest.raises(ValueError, match="version"):
        ProfileStore.open(path)


def test_surface_size_is_enforced(tmp_path):
    with ProfileStore.create(os.path.join(tmp_path, "s"), num_surf_points=31, modes=("surf",)) as st:
        with pytest.raises(ValueError, match="points"):
            st.add(_profile(0))


@pytest.mark.parametrize("modes,keys", [
    (("vol",), {"ids", "atom_off", "atom_pos"}),
    (("surf_esp",), {"ids", "atom_off", "atom_pos", "surf_pos", "surf_esp"}),
    (("vol_esp",), {"ids", "atom_off", "atom_pos", "charges"}),
    (("pharm",), {"ids", "atom_off", "atom_pos", "pharm_off", "pharm_types", "pharm_ancs", "pharm_vecs"}),
    (("vol_and_surf_esp",), {"ids", "atom_off", "atom_pos", "surf_pos", "surf_esp", "charges", "all_off",
                             "nonH", "radii", "cwh"}),
])
def test_on_disk_keys_and_dtypes(tmp_path, modes, keys):
    """The shard keys and dtypes each schema writes (the reader contract at screen.py:458-509)."""
    path = os.path.join(tmp_path, "lib.fss")
    with ProfileStore.create(path, num_surf_points=32, modes=modes, pre_centered=False) as st:
        st.add(_profile(0, with_h="vol_and_surf_esp" in modes))
        st.add(_profile(1, with_h="vol_and_surf_esp" in modes))
    _, a = ProfileStore.open(path).read_shard(0)
    assert set(a) == keys
    for k, v in a.items():
        if k.endswith("_off") or k == "nonH":
            assert v.dtype == np.int64
        elif k == "pharm_types":
            assert v.dtype == np.int32
        elif k != "ids":
            assert v.dtype == np.float16
    assert a["atom_off"].tolist() == [0, 6, 12]


def test_heavy_centres_written_only_when_they_differ(tmp_path):
    path = os.path.join(tmp_path, "lib.fss")
    with ProfileStore.create(path, num_surf_points=32, modes=("vol_esp",), dtype="float32",
                             pre_centered=False) as st:
        st.add(_profile(0))
        st.add(_profile(1, noH=True))
    _, a = ProfileStore.open(path).read_shard(0)
    assert {"heavy_off", "xyz_noH"} <= set(a)


@pytest.mark.parametrize("mode", ["vol", "vol_esp", "surf", "surf_esp", "pharm", "vol_color"])
@pytest.mark.parametrize("with_h", [False, True])
def test_loader_rows_match_per_molecule_reconstruction(tmp_path, mode, with_h):
    """The vectorised float4 rows the loader uploads == the per-molecule arrays of the mode
    (ESP / heavy charge in w), including the with-H gather of heavy charges."""
    path = os.path.join(tmp_path, "lib.fss")
    modes = ALL_MODES if with_h else ("vol", "vol_esp", "surf", "surf_esp", "pharm", "vol_color")
    profs = [_profile(i, with_h=with_h, noH=(i % 2 == 1)) for i in range(5)]
    with ProfileStore.create(path, num_surf_points=32, modes=modes, dtype="float32",
                             pre_centered=False) as st:
        for p in profs:
            st.add(p)
    st = ProfileStore.open(path)
    _, a = st.read_shard(0)
    shape, pharm = _shard_channels(a, mode, st.schema)
    if mode in ("vol", "vol_color"):
        xyz = np.concatenate([p.atom_pos for p in profs])
        w = np.zeros(len(xyz), np.float32)
    elif mode == "vol_esp":
        xyz = np.concatenate([p.atom_pos_noH if p.atom_pos_noH is not None else p.atom_pos for p in profs])
        w = np.concatenate([p.partial_charges[p.nonH_atoms_idx] if with_h else p.partial_charges
                            for p in profs])
    elif mode in ("surf", "surf_esp"):
        xyz = np.concatenate([p.surf_pos for p in profs])
        w = np.concatenate([p.surf_esp for p in profs]) if mode == "surf_esp" else np.zeros(len(xyz), np.float32)
    if mode != "pharm":
        rows, counts = shape
        np.testing.assert_array_equal(rows[:, :3], xyz)
        np.testing.assert_array_equal(rows[:, 3], w)
        assert counts.sum() == len(rows) and len(counts) == len(profs)
    if mode in ("pharm", "vol_color"):
        anc, vec, typ, pc = pharm
        np.testing.assert_array_equal(anc[:, :3], np.concatenate([p.pharm_ancs for p in profs]))
        np.testing.assert_array_equal(vec[:, :3], np.concatenate([p.pharm_vecs for p in profs]))
        np.testing.assert_array_equal(typ, np.concatenate([p.pharm_types for p in profs]))
        assert pc.tolist() == [len(p.pharm_types) for p in profs]


def test_schema_rules():
    s = schema_from_modes(["esp"])                 # legacy alias of surf_esp
    assert s["surf"] and s["surf_esp"] and not s["charges"] and not s["pharm"]
    assert store_supports(schema_from_modes(["vol_color"]), "pharm")
    assert not store_supports(schema_from_modes(["vol"]), "vol_esp")
    with pytest.raises(ValueError):
        schema_from_modes(["bogus"])


def test_loader_shaep_rows_match_profiles(tmp_path):
    """vol_and_surf_esp loads heavy atoms as the shape rows and the with-H ShaEP channel:
    surface rows (x, y, z, ESP), atom rows (x, y, z, partial charge), radii and counts."""
    from fast_shepherd_score_amd.store import _shard_shaep
    path = os.path.join(tmp_path, "lib.fss")
    profs = [_profile(i, with_h=True, total=7 + i) for i in range(5)]
    with ProfileStore.create(path, num_surf_points=32, modes=("vol_and_surf_esp",), dtype="float32",
                             pre_centered=False) as st:
        for p in profs:
            st.add(p)
    st = ProfileStore.open(path)
    _, a = st.read_shard(0)
    (rows, counts), pharm = _shard_channels(a, "vol_and_surf_esp", st.schema)
    assert pharm is None
    np.testing.assert_array_equal(rows[:, :3], np.concatenate([p.atom_pos for p in profs]))
    assert not rows[:, 3].any() and counts.tolist() == [len(p.atom_pos) for p in profs]
    srf, sc, atm, rad, ac = _shard_shaep(a)
    np.testing.assert_array_equal(srf[:, :3], np.concatenate([p.surf_pos for p in profs]))
    np.testing.assert_array_equal(srf[:, 3], np.concatenate([p.surf_esp for p in profs]))
    assert sc.tolist() == [32] * 5
    np.testing.assert_array_equal(atm[:, :3], np.concatenate([p.centers_w_H for p in profs]))
    np.testing.assert_array_equal(atm[:, 3], np.concatenate([p.partial_charges for p in profs]))
    np.testing.assert_array_equal(rad, np.concatenate([p.radii for p in profs]))
    assert ac.tolist() == [7 + i for i in range(5)]


def test_centring_on_load_matches_a_pre_centred_store(tmp_path):
    """load_library(center=True) of a store written without centring moves every channel of each
    conformer by its atom mean, as a pre-centred store does at write time."""
    from fast_shepherd_score_amd.store import _center_shard
    profs = [_profile(s, with_h=True) for s in range(4)]
    arrays = {}
    for pre in (False, True):
        path = os.path.join(tmp_path, f"lib{int(pre)}.fss")
        with ProfileStore.create(path, num_surf_points=32, modes=ALL_MODES, dtype="float32",
                                 pre_centered=pre) as st:
            for p in profs:
                st.add(p)
        arrays[pre] = ProfileStore.open(path).read_shard(0)[1]
    c = _center_shard(arrays[False])
    for k in ("atom_pos", "surf_pos", "pharm_ancs", "cwh"):
        np.testing.assert_allclose(c[k], arrays[True][k], atol=1e-5, err_msg=k)
    np.testing.assert_array_equal(c["pharm_vecs"], arrays[True]["pharm_vecs"])


def test_reference_store_helpers(tmp_path):
    """iter_shards / read_profiles / add_profile / center_to / from_molecule(modes=...) as in
    screen.py:178-197, 420-424, 556-565."""
    path = os.path.join(tmp_path, "lib.fss")
    profs = [_profile(s) for s in range(5)]
    with ProfileStore.create(path, num_surf_points=32, modes=("vol", "pharm"), dtype="float32",
                             shard_size=2, pre_centered=False) as st:
        for p in profs:
            st.add_profile(p)
        with pytest.raises(TypeError):
            st.add_profile(object())
    store = ProfileStore.open(path)
    shards = list(store.iter_shards())
    assert [len(s) for s in shards] == [2, 2, 1]
    again = store.read_profiles(1)
    np.testing.assert_array_equal(again[0].atom_pos, shards[1][0].atom_pos)
    assert [p.id for s in shards for p in s] == [100 + s for s in range(5)]
    q = MoleculeProfile.from_molecule(profs[0], modes=("pharm",))
    assert q.surf_pos is None and q.partial_charges is None and q.pharm_types is not None
    mu = q.atom_pos.mean(0)
    anc = q.pharm_ancs.copy()
    q.center_to(mu)
    np.testing.assert_allclose(q.atom_pos.mean(0), 0.0, atol=1e-5)
    np.testing.assert_allclose(q.pharm_ancs, anc - mu, atol=1e-6)
    np.testing.assert_array_equal(q.pharm_vecs, profs[0].pharm_vecs)
