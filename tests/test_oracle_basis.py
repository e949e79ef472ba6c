"""
Pins the oracle's basis routines (and the product's host-side harmonics) against
the golden fixtures regenerated from the reference's own basis tests, closed forms
and the addition theorem.  CPU only.
"""

import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name))


def test_local_eval_matches_golden():
    # tests/host/test_c_sph_harm.py:45-78 : C kernel vs lpmv form at 1e-14 absolute
    g = _load("sph_harm_L20.npz")
    L = int(g["L"])
    for i in range(len(g["charges"])):
        e = O.local_eval(L, g["moments"], g["positions"][i] - g["centre"]) * g["charges"][i]
        assert abs(e - g["energy"][i]) < 1e-14 * max(1.0, abs(g["energy"][i])) * 10


def test_dot_vec_matches_golden():
    # tests/host/test_c_sph_harm.py:129-165 : dot_vec . L == phi (reference bar 1e-13
    # relative).  The golden phi is exact (40-digit), so the bound is stated against the
    # sum of |terms|: one case cancels 100-fold and sits at 1.4e-13 relative.
    g = _load("dot_vec_L20.npz")
    L = int(g["L"])
    rels = []
    for e, p, ph in zip(g["expansions"], g["points"], g["phi"]):
        v = O.dot_vec(L, 1.0, p)
        assert abs(np.dot(v, e) - ph) <= 2e-15 * np.dot(np.abs(v), np.abs(e))
        rels.append(abs(np.dot(v, e) - ph) / abs(ph))
    assert np.median(rels) < 1e-14


def test_multipole_matches_golden():
    # tests/host/test_pbc_1.py:38-62 : multipole_exp vs lpmv form, L=26, 1e-12
    g = _load("multipole_L26.npz")
    L = int(g["L"])
    for r, th, ph, ref in zip(g["radii"], g["thetas"], g["phis"], g["multipole"]):
        disp = np.array([r * math.sin(th) * math.cos(ph), r * math.sin(th) * math.sin(ph), r * math.cos(th)])
        v = O.p2m(L, 1.0, disp)
        rel = np.maximum(np.abs(ref), 1.0)
        assert np.max(np.abs(v - ref) / rel) < 1e-12


def test_p2m_equals_dot_vec():
    # tests/host/test_c_sph_harm.py:168-200 : the two vectors agree (M == E)
    rng = np.random.RandomState(9476213)
    for _ in range(5):
        p = rng.uniform(size=3)
        assert np.max(np.abs(O.p2m(20, 1.0, p) - O.dot_vec(20, 1.0, p))) < 1e-15


def test_local_exp_matches_golden():
    # tests/host/test_c_sph_harm.py:172-188 : create_local_exp vs closed form
    g = _load("local_exp_L12.npz")
    L = int(g["L"])
    for p, ref in zip(g["points"], g["local"]):
        v = O.p2l(L, 1.0, p)
        assert np.max(np.abs(v - ref)) < 1e-10


def test_addition_theorem():
    # P2L at s then local evaluation at x reproduces 1/|x - s| to truncation (|x|/|s|)^L
    rng = np.random.default_rng(3)
    for _ in range(10):
        s = rng.uniform(-1, 1, 3)
        s *= 2.0 / np.linalg.norm(s)
        x = rng.uniform(-1, 1, 3)
        x *= 0.2 / np.linalg.norm(x)
        lexp = O.p2l(20, 1.0, s)
        v = O.local_eval(20, lexp, x)
        assert abs(v - 1.0 / np.linalg.norm(x - s)) < 1e-14 * 10 + (0.1 ** 20) * 10


def test_addition_theorem_fails_for_wrong_sign():
    """A planted mistake (conjugating the wrong operand) must fail the same check."""
    s = np.array([0.3, 1.9, 0.2])
    x = np.array([0.1, 0.5, -0.1])
    lexp = O.p2l(12, 1.0, s)
    L = 12
    bad = lexp.copy()
    bad[L * L:] *= -1.0          # Y_l^{+m} instead of Y_l^{-m}
    good = O.local_eval(L, lexp, x)
    wrong = O.local_eval(L, bad, x)
    exact = 1.0 / np.linalg.norm(x - s)
    assert abs(good - exact) < 1e-5
    assert abs(wrong - exact) > 1e-3


def test_product_harmonics_match_closed_form():
    # coulomb_kmc_amd.harmonics (used by the lattice operator and mirrored by the kernels)
    from scipy.special import lpmv
    from coulomb_kmc_amd.harmonics import regular_columns, irregular_columns
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1, 1, (20, 3))
    R = regular_columns(15, pts)
    I = irregular_columns(15, pts)
    for p, x in enumerate(pts):
        r = np.linalg.norm(x)
        th = math.atan2(math.hypot(x[0], x[1]), x[2])
        ph = math.atan2(x[1], x[0])
        for l in range(16):
            for m in range(l + 1):
                y = math.sqrt(math.factorial(l - m) / math.factorial(l + m)) * lpmv(m, l, math.cos(th)) \
                    * complex(math.cos(m * ph), math.sin(m * ph))
                assert abs(R[p, l, m] - r ** l * y) < 1e-13 * max(1.0, abs(r ** l * y))
                assert abs(I[p, l, m] - y / r ** (l + 1)) < 1e-12 * max(1.0, abs(y / r ** (l + 1)))


def test_half_layout_roundtrip():
    from coulomb_kmc_amd.harmonics import expansion_full_to_half, expansion_half_to_full
    L = 9
    full = O.p2l(L, 0.7, np.array([1.1, -0.3, 0.4]))
    back = expansion_half_to_full(expansion_full_to_half(full, L), L)
    assert np.max(np.abs(back - full)) < 1e-15 * np.max(np.abs(full)) * 10
