"""Scoring constants of the ShEPhERD similarity functions.

The values are data the reference defines and every implementation must share, restated from
``shepherd_score/score/constants.py:9-36`` (shape alpha spline, Coulomb / lambda scaling,
pharmacophore types, per-type Gaussian widths, directional set) and the type categories of
``shepherd_score/score/analytical_gradients/_torch.py:20-25, 176-211``.
"""
from __future__ import annotations

import math

import numpy as np

# Surface-point count -> shape alpha (quadratic spline through these knots).
_NPOINTS = np.array([50, 100, 150, 200, 300, 400], dtype=np.float64)
_NPOINTS_ALPHAS = np.array([0.6011, 0.8668, 1.022, 1.118, 1.216, 1.258], dtype=np.float64)


def ALPHA(n_points):
    """Shape Gaussian width for a surface of ``n_points`` points (constants.py:13)."""
    from scipy.interpolate import interp1d
    return float(interp1d(_NPOINTS, _NPOINTS_ALPHAS, "quadratic")(n_points))


COULOMB_SCALING = 1e4 / (4 * 55.263 * np.pi)   # eV*A/e^2
LAM_SCALING = COULOMB_SCALING ** 2

P_TYPES = ("Acceptor", "Donor", "Aromatic", "Hydrophobe", "Halogen", "Cation", "Anion",
           "ZnBinder", "Dummy")
P_TYPES_LWRCASE = tuple(p.lower() for p in P_TYPES)
P_ALPHAS = {"acceptor": 1.0, "donor": 1.0, "aromatic": 0.7, "hydrophobe": 0.7, "cation": 1.0,
            "anion": 1.0, "znbinder": 1.0, "halogen": 1.0, "dummy": 1.0}
DIRECTIONAL_PTYPES = frozenset({"acceptor", "donor", "aromatic", "halogen"})

_NONDIRECTIONAL = {"hydrophobe", "znbinder", "anion", "cation"}
_DIRECTIONAL = {"acceptor", "donor", "halogen"}
_AROMATIC = {"aromatic"}

# Tversky sigma per similarity name (pharmacophore_scoring.py:503-512).
SIGMA_MAP = {"tversky": 0.95, "tversky_ref": 1.0, "tversky_fit": 0.05}


def gaussian_k(alpha: float) -> float:
    """Normaliser K = (pi / (2 alpha))^1.5 of a Gaussian pair term."""
    return (math.pi / (2.0 * alpha)) ** 1.5


def pharm_tables(directionless: bool = False, extended_points: bool = False,
                 only_extended: bool = False):
    """Per-type (alpha, K, category) lists in P_TYPES order.

    Category 0 = isotropic point Gaussian (w = 1), 1 = directional (cosine weight clamped to
    [0, 1]), 2 = aromatic (|cosine| weight), 3 = ignored (Dummy).  ``directionless`` maps every
    real type to category 0 (ROCS-style colour), keeping the per-type alpha, and turns extended
    points off (pharmacophore_scoring.py:517-523).  ``extended_points`` scores the directional
    types by the overlap of their extended points anchor + unit vector, plus the plain anchor
    overlap (category 4) or without it when ``only_extended`` (category 5)
    (pharmacophore_scoring.py:284-370, analytical_gradients/_torch.py:288-357).
    """
    alphas, ks, cats = [], [], []
    for p in P_TYPES_LWRCASE:
        a = P_ALPHAS.get(p, 1.0)
        alphas.append(a)
        ks.append(gaussian_k(a))
        if p == "dummy":
            cats.append(3)
        elif directionless:
            cats.append(0)
        elif p in _NONDIRECTIONAL:
            cats.append(0)
        elif p in _DIRECTIONAL:
            cats.append((5 if only_extended else 4) if extended_points else 1)
        elif p in _AROMATIC:
            cats.append(2)
        else:
            cats.append(3)
    return alphas, ks, cats
