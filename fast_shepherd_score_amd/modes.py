"""Alignment mode registry: result attributes and per-mode defaults.

The table contents are the reference's public contract (shepherd_score/accel/_modes.py:18-61):
the in-place result attribute names written by ``MoleculePairBatch.align_with_<mode>``, the
legacy aliases, and the per-mode default restart / step counts used when ``num_repeats`` /
``max_num_steps`` are None.
"""
MODE_ATTRS = {
    "vol": ("transform_vol_noH", "sim_aligned_vol_noH"),
    "vol_esp": ("transform_vol_esp_noH", "sim_aligned_vol_esp_noH"),
    "surf": ("transform_surf", "sim_aligned_surf"),
    "surf_esp": ("transform_surf_esp", "sim_aligned_surf_esp"),
    "vol_and_surf_esp": ("transform_vol_and_surf_esp", "sim_aligned_vol_and_surf_esp"),
    "pharm": ("transform_pharm", "sim_aligned_pharm"),
    "vol_color": ("transform_vol_color", "sim_aligned_vol_color"),
}
CANONICAL_MODES = tuple(MODE_ATTRS)
LEGACY_MODE_ALIASES = {"esp": "surf_esp", "esp_combo": "vol_and_surf_esp"}
MODE_SEEDS = {"vol": 10, "surf": 8, "surf_esp": 8, "vol_esp": 16, "vol_and_surf_esp": 8,
              "pharm": 32, "vol_color": 16}
MODE_STEPS = {"vol": 30, "surf": 40, "surf_esp": 40, "vol_esp": 50, "vol_and_surf_esp": 60,
              "pharm": 50, "vol_color": 40}
# Modes the HIP backend implements.
HIP_MODES = ("vol", "vol_esp", "surf", "surf_esp", "pharm", "vol_color", "vol_and_surf_esp")


def canonical(mode: str) -> str:
    return LEGACY_MODE_ALIASES.get(mode, mode)
