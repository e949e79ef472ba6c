"""
Pins the oracle's extended-precision checker (Oracle.propose_check), the reference the
GPU paths' proposal energies are measured against:

* on small systems, where the reference-ordered FP64 propose loses little to rounding,
  checker (long double) == reference-ordered propose to 1e-12 in every boundary
  condition, component by component (near, far, self + long range);
* the long-range term in the checker's expanded difference form equals the difference
  of two whole-system long-range energies (the reference's form,
  coulomb_kmc/kmc_full_long_range.py:843-858) computed in the checker's own
  precision (a plausible slip - a dropped linear term, a sign - fails it);
* the checker on double and on long double agree to FP64 rounding, and on a system
  whose whole-system long-range energy is large the reference ordering drifts from the
  extended value by more than the checker's own double form does.
"""

import numpy as np
import pytest

import oracle as O
from helpers import neutral_system, ball_moves

from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout


def _oracle(pos, q, E, bc, L=12, r=3, max_move=1.0):
    lr = lr_operator_reference_layout(L, E) if bc in ("pbc", "ff-only") else None
    o = O.Oracle(pos, q, E, r, L, bc=bc, max_move=max_move, lr=lr)
    o.initialise()
    return o


def _moves(rng, pos, N, P, k, radius, E, bc):
    pids = rng.choice(N, P, replace=False)
    mv = ball_moves(rng, pos, pids, k, radius, E, bc)
    es = np.arange(P + 1) * k
    return pids, es, np.concatenate([m[1] for m in mv])


@pytest.mark.parametrize("bc", ("free_space", "27", "pbc", "ff-only"))
def test_checker_equals_reference_order_small(bc):
    rng = np.random.default_rng(11)
    N, E = 200, 5.0
    pos, q = neutral_system(rng, N, E)
    o = _oracle(pos, q, E, bc)
    pids, es, npos = _moves(rng, pos, N, 12, 4, 1.0, E, bc)
    ref, rc = o.propose_arrays(pids, es, npos, comps=True)
    ext, ec = o.propose_check(pids, es, npos, extended=True, comps=True)
    assert np.max(np.abs(ext - ref)) < 1e-12
    # component by component: near, far (new and old), self + long range
    assert np.max(np.abs(ec - rc)) < 1e-12
    d64 = o.propose_check(pids, es, npos, extended=False)
    assert np.max(np.abs(d64 - ext)) < 1e-13


def test_checker_new_charges():
    rng = np.random.default_rng(12)
    N, E = 160, 4.0
    pos, q = neutral_system(rng, N, E)
    o = _oracle(pos, q, E, "pbc")
    pids, es, npos = _moves(rng, pos, N, 6, 3, 1.0, E, "pbc")
    nq = rng.choice([-2.0, 0.5, 1.0], es[-1])
    ref = o.propose_arrays(pids, es, npos, new_q=nq)
    ext = o.propose_check(pids, es, npos, new_q=nq, extended=True)
    assert np.max(np.abs(ext - ref)) < 1e-12


def test_long_range_difference_form_matches_whole_system_difference():
    """ff-only isolates the long-range term: dE = E_lr(M', E') - E_lr(M, E) (eu = 1).
    The whole-system energies are formed with the oracle's own P2M / dot-vector and
    bilinear form (p2m, dot_vec, lr operator) and subtracted in numpy FP64 on a
    small system whose energies are O(1), so their difference is exact to ~1e-14."""
    rng = np.random.default_rng(13)
    N, E, L = 40, 3.0, 10
    pos, q = neutral_system(rng, N, E)
    lr = lr_operator_reference_layout(L, E)
    o = O.Oracle(pos, q, E, 2, L, bc="ff-only", max_move=1.0, lr=lr)
    o.initialise()
    data, indptr, indices, dip = lr
    L2 = L * L

    def form(Mv, Ev):
        Rm = np.zeros(L2)
        Ri = np.zeros(L2)
        for row in range(L2):
            for k in range(indptr[row], indptr[row + 1]):
                Rm[row] += data[k] * Mv[indices[k]]
                Ri[row] += data[k] * Mv[indices[k] + L2]
        dot = np.dot(Rm, Ev[:L2]) + np.dot(Ri, Ev[L2:])
        d = (dip[0] * Mv[3] * Ev[3] + dip[0] * Mv[3] * Ev[1] - dip[1] * Mv[L2 + 3] * Ev[L2 + 3]
             + dip[1] * Mv[L2 + 3] * Ev[L2 + 1] + dip[2] * Mv[2] * Ev[2])
        return 0.5 * dot + 0.5 * d

    M0, E0 = o.get_ME()
    e0 = form(M0, E0)
    for pid in (3, 17, 29):
        newp = pos[pid] + rng.uniform(-0.5, 0.5, 3)
        newp = np.where(newp < -0.5 * E, newp + E, newp)
        newp = np.where(newp >= 0.5 * E, newp - E, newp)
        M1 = M0 - O.p2m(L, q[pid], pos[pid]) + O.p2m(L, q[pid], newp)
        E1 = E0 - O.dot_vec(L, q[pid], pos[pid]) + O.dot_vec(L, q[pid], newp)
        want = form(M1, E1) - e0
        got = o.propose_check([pid], [0, 1], newp[None, :], extended=True)[0]
        assert abs(got - want) < 1e-12 * max(1.0, abs(e0))
        # a dropped linear term or a flipped sign would be off by O(|dE|)
        assert abs(want) > 1e-4


def test_reference_order_drifts_where_checker_does_not():
    """A dipolar system (charges +1 on one side, -1 on the other) in a large box: the
    whole-system long-range energy is large, so the reference's subtraction of two of
    them loses digits; the checker's double form stays at the extended value."""
    rng = np.random.default_rng(14)
    N, E, L = 400, 40.0, 12
    pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
    q = np.where(pos[:, 0] > 0, 1.0, -1.0)
    q -= q.mean()
    o = _oracle(pos, q, E, "pbc", L=L, r=3, max_move=1.5)
    pids, es, npos = _moves(rng, pos, N, 20, 4, 1.5, E, "pbc")
    ref = o.propose_arrays(pids, es, npos)
    ext = o.propose_check(pids, es, npos, extended=True)
    d64 = o.propose_check(pids, es, npos, extended=False)
    err_ref = np.max(np.abs(ref - ext))
    err_d64 = np.max(np.abs(d64 - ext))
    assert err_d64 < 1e-13
    assert err_d64 <= err_ref
    assert abs(o.lr_energy()) > 1.0


def test_de_error_measure():
    ref = np.array([1.0, -2e-4, 0.5, 0.0])
    a = ref + np.array([1e-12, 1e-14, -5e-13, 2e-15])
    st = O.de_error(a, ref)
    # floored relative: |d| / max(|ref|, 1e-3)
    np.testing.assert_allclose(st["max_rel"], max(1e-12, 1e-14 / 1e-3, 5e-13 / 0.5, 2e-15 / 1e-3), rtol=1e-3)
    assert st["below_floor"] == 2 and st["proposals"] == 4
