"""
Pins the oracle's inject / extract pieces (coulomb_kmc/kmc_inject_extract.py:82-237)
on the CPU against independent physics, following the reference's own tests:

* propose_extract against the direct energy difference of the remaining charges, on
  the reference's inputs (tests/host/test_extract.py:49-150: N=20, extent 2*3.1416,
  RandomState(3251), r=3, l=12; relative error < 3.1e-4 on every draw, duplicated
  rows identical to 1e-12);
* propose_inject against the direct energy difference of the grown system
  (tests/host/test_inject.py:50-160: same inputs, relative error < 2e-4);
* compute_energy (the group energy) against brute-force pair sums, 27-image sums and
  an independent Ewald sum of the group alone;
* per-charge energies (get_energy) against brute-force q_i sum_j q_j / r_ij in free
  space.
"""

import numpy as np
import pytest

import oracle as O
from ewald import ewald_energy, free_space_energy, nearest27_energy
from helpers import neutral_system

from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout

REFS = {
    "free_space": lambda p, q, E: free_space_energy(p, q),
    "27": lambda p, q, E: nearest27_energy(p, q, E),
    "pbc": lambda p, q, E: ewald_energy(p, q, E),
}


def _oracle(pos, q, E, bc, L=12, r=3, max_move=1.0):
    lr = lr_operator_reference_layout(L, E) if bc in ("pbc", "ff-only") else None
    o = O.Oracle(pos, q, E, r, L, bc=bc, max_move=max_move, lr=lr)
    o.energy = o.initialise()
    return o


def _reference_inputs():
    # tests/host/test_extract.py:66-75
    N, E = 20, 2 * 3.1416
    rng = np.random.RandomState(3251)
    pi = rng.uniform(low=-0.5 * E, high=0.5 * E, size=(N, 3))
    qi = np.array([(-1.0) ** px for px in range(N)])
    return N, E, rng, pi, qi


def _draw_pairs(rng, N, q, num):
    # tests/host/test_extract.py:116-137: num positive then num negative distinct ids
    inds, available = [], set(range(N))
    for sign in (1, -1):
        for _ in range(num):
            ind = rng.randint(0, N)
            while (sign * q[ind] < 0) or (ind not in available):
                ind = rng.randint(0, N)
            inds.append(ind)
            available.remove(ind)
    return inds


@pytest.mark.parametrize("bc", ("free_space", "27", "pbc"))
def test_propose_extract_against_direct(bc):
    N, E, rng, pi, qi = _reference_inputs()
    o = _oracle(pi, qi, E, bc)
    phi0 = REFS[bc](pi, qi, E)
    assert abs(o.energy - phi0) / abs(phi0) < 1e-5
    for _ in range(20):
        inds = _draw_pairs(rng, N, qi, rng.randint(1, 10))
        d = o.propose_extract(inds)[0]
        keep = np.array(sorted(set(range(N)) - set(inds)))
        diff = REFS[bc](pi[keep], qi[keep], E) - phi0
        err = abs(d - diff) / abs(diff)
        assert err < 3.1e-4 or abs(d - diff) < 1e-5, (err, d, diff)
        d2 = o.propose_extract((inds, inds))
        assert abs(d2[0] - d) < 1e-12 and abs(d2[1] - d) < 1e-12


@pytest.mark.parametrize("bc", ("free_space", "27", "pbc"))
def test_propose_inject_against_direct(bc):
    N, E, rng, pi, qi = _reference_inputs()
    phi_full = REFS[bc](pi, qi, E)
    for _ in range(6):
        inds = _draw_pairs(rng, N, qi, rng.randint(1, 10))
        keep = np.array(sorted(set(range(N)) - set(inds)))
        o = _oracle(pi[keep], qi[keep], E, bc)
        phi0 = REFS[bc](pi[keep], qi[keep], E)
        assert abs(o.energy - phi0) / abs(phi0) < 1e-5
        d = o.propose_inject(pi[inds], qi[inds])
        diff = phi_full - phi0
        assert abs(d - diff) / abs(diff) < 2e-4, (d, diff)


@pytest.mark.parametrize("bc", ("free_space", "27", "pbc"))
def test_group_energy_against_direct_sums(bc):
    E = 5.0
    rng = np.random.RandomState(77)
    pos = rng.uniform(-0.5 * E, 0.5 * E, (40, 3))
    q = np.array([(-1.0) ** i for i in range(40)])
    o = _oracle(pos, q, E, bc, L=16)
    for m in (2, 4, 8):
        sel = rng.choice(40, m, replace=False)
        g, gq = pos[sel], q[sel]
        if abs(gq.sum()) > 0:
            gq = gq - gq.mean()
        ref = REFS[bc](g, gq, E)
        tol = 1e-12 if bc != "pbc" else 2e-6
        assert abs(o.group_energy(g, gq) - ref) <= tol * max(abs(ref), 1.0), (m, o.group_energy(g, gq), ref)


def test_charge_energy_free_space_brute_force():
    E = 6.0
    rng = np.random.RandomState(5)
    N = 200
    pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
    q = np.where(rng.uniform(size=N) < 0.5, -1.0, 1.0)
    o = _oracle(pos, q, E, "free_space", L=12, r=3)
    ids = rng.choice(N, 25, replace=False)
    ce = o.charge_energy(ids)
    for k, i in enumerate(ids):
        d = np.linalg.norm(pos - pos[i], axis=1)
        d[i] = np.inf
        ref = q[i] * np.sum(q / d)
        assert abs(ce[k] - ref) <= 1e-5 * max(abs(ref), 1.0), (i, ce[k], ref)
    # a charge's energy equals the energy change of removing it, sign reversed
    e_all = free_space_energy(pos, q)
    i = int(ids[0])
    keep = np.arange(N) != i
    assert abs(-ce[0] - (free_space_energy(pos[keep], q[keep]) - e_all)) <= 1e-5 * abs(ce[0])


@pytest.mark.parametrize("bc", ["free_space", "27", "pbc", "ff-only"])
def test_half_sum_of_charge_energies_is_the_energy(bc):
    """0.5 sum_i U_i = E for the oracle's per-charge energies (the identity the
    per-charge potential of initialise is checked by); free space also against the
    brute-force potential within the expansion truncation."""
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    rng = np.random.default_rng(21)
    N, E = 300, 6.0
    pos, q = neutral_system(rng, N, E)
    lr = lr_operator_reference_layout(12, E) if bc in ("pbc", "ff-only") else None
    o = O.Oracle(pos, q, E, 4, 12, bc=bc, max_move=1.0, lr=lr)
    e = o.initialise()
    u = o.charge_energy(np.arange(N))
    assert abs(0.5 * u.sum() - e) <= 1e-13 * abs(e)
    if bc == "free_space":
        d = pos[:, None, :] - pos[None, :, :]
        r = np.sqrt((d * d).sum(-1))
        np.fill_diagonal(r, np.inf)
        phi = (q[None, :] / r).sum(axis=1)
        assert np.max(np.abs(u / q - phi)) <= 1e-5 * np.max(np.abs(phi))
