"""Hierarchical initial solve (init_method='fmm') against the exact direct P2L
initialiser (init_method='direct') on the same charges.

The reference takes its initial leaf expansions from ppmd's FMM
(coulomb_kmc/kmc_fmm.py:463-476, 956-959), whose coefficients cannot be reproduced
here; what pins this path is (a) exactness where the hierarchy is not used (free space
with 8^3 leaves: only the direct near-region P2L runs), (b) agreement with the exact
initialiser to the truncation error of the translations, and (c) that this error
falls with the expansion order as an FMM's must (a wrong translation term would not
converge)."""

import numpy as np
import pytest

from helpers import neutral_system

pytestmark = pytest.mark.gpu


def _init_pair(N, E, r, L, bc, seed=3):
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(seed)
    pos, q = neutral_system(rng, N, E)
    out = {}
    for m in ("direct", "fmm"):
        k = KMCFMM(pos.copy(), q, E, r=r, l=L, boundary_condition=bc, init_method=m)
        k.initialise()
        out[m] = (k.get_expansions(), k.energy)
        k.free()
    return out


def _order_errors(a, b, L):
    L2 = L * L
    errs = []
    for l in range(L):
        idx = [l * l + l + m for m in range(-l, l + 1)]
        idx = idx + [L2 + i for i in idx]
        errs.append(np.abs(a[:, idx] - b[:, idx]).max() / np.abs(b[:, idx]).max())
    return np.array(errs)


def test_fmm_free_space_8_leaves_is_exact():
    out = _init_pair(4000, 20.0, 4, 12, "free_space")
    a, b = out["fmm"][0], out["direct"][0]
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    assert abs(out["fmm"][1] - out["direct"][1]) <= 1e-12 * abs(out["direct"][1])


@pytest.mark.parametrize("bc", ["pbc", "free_space", "27"])
def test_fmm_matches_direct(bc):
    out = _init_pair(20000, 40.0, 5, 12, bc)
    errs = _order_errors(out["fmm"][0], out["direct"][0], 12)
    assert errs.max() < 2e-4, errs
    e_f, e_d = out["fmm"][1], out["direct"][1]
    assert abs(e_f - e_d) <= 1e-6 * abs(e_d)


def test_fmm_error_falls_with_order():
    # translations run at order L + 4: 8 orders apart the truncation error of the
    # well-separated interactions (ratio ~0.76 per order) falls by ~9x
    o4 = _init_pair(8000, 30.0, 5, 4, "pbc")
    o12 = _init_pair(8000, 30.0, 5, 12, "pbc")
    e4 = _order_errors(o4["fmm"][0], o4["direct"][0], 4)
    e12 = _order_errors(o12["fmm"][0], o12["direct"][0], 12)
    assert e12[:4].max() < 0.3 * e4.max(), (e4, e12)


def test_accept_after_fmm_init_matches_sweep_of_direct():
    """An accept moves the same charge in both initialisations: the expansion change
    (new - old) is the exact sweep in both, so the differences before and after agree."""
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(11)
    pos, q = neutral_system(rng, 6000, 30.0)
    ks = {}
    for m in ("direct", "fmm"):
        k = KMCFMM(pos.copy(), q, 30.0, r=5, l=10, boundary_condition="pbc", init_method=m, max_move=2.0)
        k.initialise()
        before = k.get_expansions()
        npos = pos[17] + np.array([0.9, -0.4, 0.3])
        npos = (npos + 15.0) % 30.0 - 15.0
        k.accept((17, npos))
        ks[m] = k.get_expansions() - before
        k.free()
    d = np.abs(ks["fmm"] - ks["direct"]).max()
    assert d <= 1e-12 * np.abs(ks["direct"]).max()
