"""Launch planning of the persistent optimiser for clouds of any size (host-only, no GPU).

fss_align_plan reports the launch fss_align would make.  Clouds that do not fit the workgroup's
LDS (160 KiB on gfx950) in the requested layout fall back to the expanded form (MFMA layout
3 -> 2) and, in the shape modes, to the expanded form over clouds streamed through LDS tiles
(variant 4); the other modes are refused with FSS_E_TOO_LARGE.  The default inner loop also
streams clouds whose staged LDS would leave a CU fewer than 8 wavefronts.
"""
import ctypes

import pytest

from fast_shepherd_score_amd import _lib

LDS = 160 * 1024


def plan(mode, max_ref, max_fit, variant=0, lpp=64, rpw=4, early_stop=0, B=8, ppw=1, tile_rows=0,
         groups=0):
    lib = _lib.load()
    pb = _lib.FssPairBatch()
    pb.n_pairs = 3
    # planning reads only the maxima and checks the arrays are present; nothing is dereferenced
    pb.ref_pts = pb.fit_pts = pb.ref_rng = pb.fit_rng = 8
    pb.ref_anc = pb.fit_anc = pb.ref_vec = pb.fit_vec = pb.ref_typ = pb.fit_typ = pb.ref_prng = pb.fit_prng = 8
    pb.max_ref, pb.max_fit = max_ref, max_fit
    pb.max_ref_ph = pb.max_fit_ph = max_ref if mode == _lib.MODE_PHARM else 8
    sp = _lib.FssScoreParams()
    sp.mode, sp.alpha, sp.lam = mode, 0.81, 0.3
    ap = _lib.FssAlignParams()
    ap.n_restarts, ap.max_steps = B, 10
    ap.lr, ap.beta1, ap.beta2, ap.eps = 0.1, 0.9, 0.999, 1e-8
    ap.early_stop, ap.es_tol, ap.es_patience = early_stop, 1e-5, 10
    ap.lanes_per_pose, ap.restarts_per_wg = lpp, rpw
    ap.variant, ap.groups_per_wg, ap.pairs_per_wg, ap.tile_rows = variant, groups, ppw, tile_rows
    info = _lib.FssAlignPlanInfo()
    rc = lib.fss_align_plan(ctypes.byref(pb), ctypes.byref(sp), ctypes.byref(ap), ctypes.byref(info))
    return rc, info, lib.fss_last_error()


def test_small_clouds_keep_the_requested_loop():
    for v in (2, 3):
        rc, info, _ = plan(_lib.MODE_SHAPE_ESP, 400, 400, variant=v)
        assert rc == 0 and info.variant == v and info.tile_ref == info.tile_fit == 0
        assert info.lds_bytes <= LDS


def test_mfma_layout_too_large_runs_the_expanded_form():
    # 3000-row ref: 188 MFMA tiles x 76 float4 = 228 KiB; the expanded form needs 108 KiB
    rc, info, _ = plan(_lib.MODE_SHAPE_ESP, 3000, 3000, variant=3)
    assert rc == 0 and info.variant == 2 and info.tile_ref == 0
    assert info.lds_bytes <= LDS


@pytest.mark.parametrize("mode", [_lib.MODE_SHAPE, _lib.MODE_SHAPE_ESP])
@pytest.mark.parametrize("variant", [0, 2, 3])
def test_clouds_beyond_lds_stream_through_tiles(mode, variant):
    rc, info, _ = plan(mode, 20000, 6000, variant=variant)
    assert rc == 0, _
    assert info.variant == 4 and info.tile_ref == 1024 and info.tile_fit == 1024
    assert info.lds_bytes <= LDS
    # a side that is smaller than a tile is staged whole
    rc, info, _ = plan(mode, 12000, 300)
    assert rc == 0 and info.variant == 4 and (info.tile_ref, info.tile_fit) == (1024, 300)


def test_default_loop_streams_clouds_that_starve_occupancy():
    # 7000 x 300 surface + ESP rows staged whole take 145 KiB: one two-wavefront workgroup per CU
    rc, info, _ = plan(_lib.MODE_SHAPE_ESP, 7000, 300, variant=0, lpp=64, rpw=2)
    assert rc == 0 and info.variant == 4 and info.tile_ref == 1024
    # the expanded form asked for by name stays staged; small clouds stay staged
    rc, info, _ = plan(_lib.MODE_SHAPE_ESP, 7000, 300, variant=2, lpp=64, rpw=2)
    assert rc == 0 and info.variant == 2 and info.tile_ref == 0
    rc, info, _ = plan(_lib.MODE_SHAPE_ESP, 400, 400, variant=0, lpp=64, rpw=2)
    assert rc == 0 and info.variant == 2


def test_streamed_form_on_request_and_tile_rows():
    rc, info, _ = plan(_lib.MODE_SHAPE, 200, 90, variant=4, tile_rows=64)
    assert rc == 0 and info.variant == 4 and (info.tile_ref, info.tile_fit) == (64, 64)
    rc, info, _ = plan(_lib.MODE_SHAPE, 50, 90, variant=4)
    assert rc == 0 and (info.tile_ref, info.tile_fit) == (50, 90)
    # early stop with every restart of the pair in one workgroup, groups running several restarts
    rc, info, _ = plan(_lib.MODE_SHAPE_ESP, 12000, 12000, variant=3, early_stop=1, B=50, rpw=50, groups=10)
    assert rc == 0 and info.variant == 4 and info.block == 640 and info.lds_bytes <= LDS


def test_lane_groups_of_several_wavefronts():
    rc, info, err = plan(_lib.MODE_SHAPE_ESP, 12000, 300, lpp=256, rpw=4)
    assert rc == 0 and info.variant == 4 and info.block == 1024, err
    rc, info, err = plan(_lib.MODE_SHAPE, 12000, 300, lpp=1024, rpw=10, early_stop=1, B=10, groups=1)
    assert rc == 0 and info.variant == 4 and info.block == 1024, err
    # only the streamed form takes them
    rc, _, err = plan(_lib.MODE_SHAPE, 300, 300, lpp=256, rpw=2)
    assert rc == _lib.E_INVALID and b"streamed" in err
    rc, _, err = plan(_lib.MODE_SHAPE, 12000, 300, lpp=96, rpw=1)
    assert rc == _lib.E_INVALID
    rc, _, err = plan(_lib.MODE_SHAPE, 12000, 300, lpp=512, rpw=4)
    assert rc == _lib.E_INVALID and b"1024 threads" in err


def test_tiles_shrink_to_fit_a_large_restart_tail():
    # 1300 restarts in one workgroup, two per lane group, with early stop: 20.8 KiB of losses and
    # 124.8 KiB of Adam slots leave 14.4 KiB for the tiles
    rc, info, err = plan(_lib.MODE_SHAPE, 50000, 50000, lpp=1, rpw=1300, B=1300, early_stop=1, groups=650)
    assert rc == 0, err
    assert info.variant == 4 and info.lds_bytes <= LDS and max(info.tile_ref, info.tile_fit) < 1024


def test_shape_pharm_streams_its_shape_clouds():
    rc, info, err = plan(_lib.MODE_SHAPE_PHARM, 20000, 3000)
    assert rc == 0 and info.variant == 4 and info.tile_ref == 1024, err
    # the pharmacophore rows (8 per side here) stay staged next to the tiles
    rc2, info2, _ = plan(_lib.MODE_SHAPE, 20000, 3000)
    assert info.lds_bytes - info2.lds_bytes == 16 * (4 + 4 * 8)


def test_refusals():
    rc, _, err = plan(_lib.MODE_PHARM, 40000, 40000)
    assert rc == _lib.E_TOO_LARGE and b"LDS" in err
    rc, _, err = plan(_lib.MODE_SHAPE, 20000, 20000, lpp=2, rpw=8, ppw=3)
    assert rc == _lib.E_TOO_LARGE and b"pairs_per_wg" in err
    rc, _, err = plan(_lib.MODE_SHAPE, 20000, 20000, variant=1)
    assert rc == _lib.E_TOO_LARGE and b"expanded form" in err
    rc, _, err = plan(_lib.MODE_SHAPE, 200, 200, tile_rows=-1)
    assert rc == _lib.E_INVALID
