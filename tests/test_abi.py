"""C ABI checks that need no GPU: the library loads, exports every declared symbol, the ctypes
structures match the C layout, and argument validation rejects bad calls before any launch."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from fast_shepherd_score_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fss.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fss_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert lib.fss_version().decode().startswith("fss_hip")


def test_struct_layout_matches_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "fss.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(FssPairBatch), sizeof(FssScoreParams), sizeof(FssAlignParams), sizeof(FssAlignOut));
  printf("%zu %zu %zu %zu\n", offsetof(FssPairBatch, max_ref), offsetof(FssScoreParams, ph_cat),
         offsetof(FssAlignParams, pair_order), offsetof(FssAlignOut, n_steps));
  printf("%zu %zu %zu %zu\n", sizeof(FssLibSpec), sizeof(FssScreenParams), sizeof(FssQuery), sizeof(FssHits));
  printf("%zu %zu %zu %zu\n", offsetof(FssLibSpec, max_pharm), offsetof(FssScreenParams, quat_table),
         offsetof(FssQuery, typ), offsetof(FssHits, latency_ms));
  printf("%zu %zu %zu %zu %zu\n", offsetof(FssLibSpec, max_atoms), offsetof(FssScreenParams, shape_from_surface),
         offsetof(FssQuery, rad), sizeof(FssLibRows), offsetof(FssLibRows, acounts));
  printf("%zu %zu %zu %zu\n", offsetof(FssQuery, all_scores), offsetof(FssHits, all_scores), offsetof(FssHits, n_all),
         offsetof(FssScreenParams, query_budget_bytes));
  printf("%zu %zu %zu %zu\n", offsetof(FssAlignParams, tile_rows), sizeof(FssAlignPlanInfo),
         offsetof(FssAlignPlanInfo, lds_bytes), offsetof(FssScreenParams, groups_per_wg));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes[:4] == [ctypes.sizeof(_lib.FssPairBatch), ctypes.sizeof(_lib.FssScoreParams),
                         ctypes.sizeof(_lib.FssAlignParams), ctypes.sizeof(_lib.FssAlignOut)]
    assert sizes[4:8] == [_lib.FssPairBatch.max_ref.offset, _lib.FssScoreParams.ph_cat.offset,
                          _lib.FssAlignParams.pair_order.offset, _lib.FssAlignOut.n_steps.offset]
    assert sizes[8:12] == [ctypes.sizeof(_lib.FssLibSpec), ctypes.sizeof(_lib.FssScreenParams),
                           ctypes.sizeof(_lib.FssQuery), ctypes.sizeof(_lib.FssHits)]
    assert sizes[12:16] == [_lib.FssLibSpec.max_pharm.offset, _lib.FssScreenParams.quat_table.offset,
                            _lib.FssQuery.typ.offset, _lib.FssHits.latency_ms.offset]
    assert sizes[16:21] == [_lib.FssLibSpec.max_atoms.offset, _lib.FssScreenParams.shape_from_surface.offset,
                            _lib.FssQuery.rad.offset, ctypes.sizeof(_lib.FssLibRows),
                            _lib.FssLibRows.acounts.offset]
    assert sizes[21:25] == [_lib.FssQuery.all_scores.offset, _lib.FssHits.all_scores.offset,
                            _lib.FssHits.n_all.offset, _lib.FssScreenParams.query_budget_bytes.offset]
    assert sizes[25:] == [_lib.FssAlignParams.tile_rows.offset, ctypes.sizeof(_lib.FssAlignPlanInfo),
                          _lib.FssAlignPlanInfo.lds_bytes.offset, _lib.FssScreenParams.groups_per_wg.offset]


def test_validation_rejects_bad_arguments_without_launching():
    """These calls fail in host-side validation, before any device work."""
    lib = _lib.load()
    sp = _lib.FssScoreParams()
    sp.mode = 7
    pb = _lib.FssPairBatch()
    rc = lib.fss_self_overlap(_lib.byref(pb), _lib.byref(sp), None, None)
    assert rc == _lib.E_INVALID and b"unknown mode" in lib.fss_last_error()
    sp.mode, sp.alpha = _lib.MODE_PHARM, 0.81
    sp.ph_cat[2] = 6          # categories are 0-5 (5: extended point only)
    rc = lib.fss_self_overlap(_lib.byref(pb), _lib.byref(sp), None, None)
    assert rc == _lib.E_INVALID and b"ph_cat[2] = 6" in lib.fss_last_error()
    sp.ph_cat[2] = 0
    sp.mode, sp.alpha = _lib.MODE_SHAPE, 0.81
    rc = lib.fss_self_overlap(_lib.byref(pb), _lib.byref(sp), None, None)
    assert rc == _lib.E_INVALID and b"NULL" in lib.fss_last_error()
    ap = _lib.FssAlignParams()
    ap.n_restarts, ap.lanes_per_pose, ap.restarts_per_wg, ap.early_stop = 50, 3, 50, 1
    out = _lib.FssAlignOut()
    pb.n_pairs = 1
    pb.ref_pts = pb.fit_pts = pb.ref_rng = pb.fit_rng = 8
    out.q = out.t = out.pose_score = out.best_score = out.best_idx = out.best_tf = out.n_steps = 8
    rc = lib.fss_align(_lib.byref(pb), _lib.byref(sp), _lib.byref(ap), ctypes.c_void_p(8), _lib.byref(out), None)
    assert rc == _lib.E_INVALID and b"power of two" in lib.fss_last_error()
    ap.lanes_per_pose, ap.restarts_per_wg = 16, 10
    rc = lib.fss_align(_lib.byref(pb), _lib.byref(sp), _lib.byref(ap), ctypes.c_void_p(8), _lib.byref(out), None)
    assert rc == _lib.E_INVALID and b"early stop" in lib.fss_last_error()
    with pytest.raises(_lib.FssError):
        _lib.call("fss_adam_step", 1, 0, 0, 0, 0, 0, 1, 0.1, 0.9, 0.999, 1e-8, 0, 1, 0)


@pytest.mark.parametrize("rem,cap", [([10, 10, 10], 12), ([0, 5, 100], 64), ([3], 1000), ([7, 1], 2),
                                      ([1000] * 7, 65536), ([5, 0, 0, 9], 3), ([], 5), ([4, 4], 0)])
def test_native_launch_planner_matches_host_restatement(rem, cap):
    """fss_plan_launch (the screening worker's planner) == screen.plan_launch, host-only."""
    from fast_shepherd_score_amd.screen import plan_launch
    lib = _lib.load()
    n = len(rem)
    r = (ctypes.c_int64 * max(n, 1))(*rem)
    take = (ctypes.c_int64 * max(n, 1))()
    assert lib.fss_plan_launch(n, r, cap, take) == 0
    assert list(take)[:n] == plan_launch(rem, cap)


def test_screening_abi_rejects_bad_arguments_without_device_work():
    lib = _lib.load()
    spec = _lib.FssLibSpec()
    out = ctypes.c_void_p()
    spec.channels, spec.max_conformers = 16, 10
    assert lib.fss_lib_create(_lib.byref(spec), ctypes.byref(out)) == _lib.E_INVALID
    assert b"channels" in lib.fss_last_error()
    spec.channels, spec.max_surface = _lib.LIB_SHAEP | _lib.LIB_SURFACE, 10
    assert lib.fss_lib_create(_lib.byref(spec), ctypes.byref(out)) == _lib.E_INVALID
    assert b"already holds the surface rows" in lib.fss_last_error()
    spec.channels, spec.max_surface = _lib.LIB_SURFACE, 0
    assert lib.fss_lib_create(_lib.byref(spec), ctypes.byref(out)) == _lib.E_INVALID
    assert b"max_surface" in lib.fss_last_error()
    spec.channels, spec.max_surface, spec.max_atoms = _lib.LIB_SHAEP, 0, 10
    assert lib.fss_lib_create(_lib.byref(spec), ctypes.byref(out)) == _lib.E_INVALID
    assert b"max_surface" in lib.fss_last_error()
    assert lib.fss_lib_append(None, 1, None, None) == _lib.E_INVALID
    spec.channels, spec.max_conformers, spec.max_points = _lib.LIB_SHAPE, 0, 10
    assert lib.fss_lib_create(_lib.byref(spec), ctypes.byref(out)) == _lib.E_INVALID
    assert lib.fss_screen_create(None, None, None, ctypes.byref(out)) == _lib.E_INVALID
    t = ctypes.c_int64()
    assert lib.fss_screen_submit(None, None, ctypes.byref(t)) == _lib.E_INVALID
    assert lib.fss_screen_poll(None, 1, 0, None) == _lib.E_INVALID
    assert lib.fss_lib_destroy(None) == 0 and lib.fss_screen_destroy(None) == 0
