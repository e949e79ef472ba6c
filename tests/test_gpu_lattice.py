"""
GPU parity of the device-resident lattice steps (Algorithm 2 of the reference's
examples: docs/source/example_bookkeeping.rst, examples/propose_with_dats_example.py:
217-450) against the same steps restated on the CPU: proposals and overlap masks
by brute force in numpy (exact), energy differences from the oracle's
propose_with_dats (1e-10 relative), rates and the selection from the running sum in
(charge, offset) order, accepts through the oracle.  The selected proposals and the
accepted positions must agree exactly over a run of steps (eager and graph-replayed).
"""

import numpy as np
import pytest

import oracle as O
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _dense_system(seed, N=400, la=3.3, Ns=6):
    """N charges +-1 on random sites of the la/3 sub-lattice of the example's box."""
    rng = np.random.default_rng(seed)
    E = la * Ns
    h = la / 3.0
    n1 = 3 * Ns
    sites = rng.choice(n1 ** 3, N, replace=False)
    ijk = np.stack([sites % n1, (sites // n1) % n1, sites // (n1 * n1)], axis=1)
    pos = -0.5 * E + 0.5 * h + h * ijk
    q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    return pos, q, E, rng


def _cpu_proposals(cur, offsets, E, wrap):
    newp = cur[:, None, :] + offsets[None, :, :]
    ok = np.ones(newp.shape[:2], bool)
    if wrap:
        newp = np.where(newp < -0.5 * E, newp + E, newp)
        newp = np.where(newp > 0.5 * E, newp - E, newp)
    else:
        ok &= np.all((newp >= -0.5 * E) & (newp <= 0.5 * E), axis=2)
    return newp, ok


def _cpu_masks(cur, newp, ok, E, tol, wrap):
    N, M = newp.shape[:2]
    d = cur[None, None, :, :] - newp[:, :, None, :]
    if wrap:
        d -= E * np.rint(d / E)
    r2 = np.sum(d * d, axis=3)
    r2[np.arange(N), :, np.arange(N)] = np.inf
    return ok & np.all(r2 >= tol, axis=2)


def _cpu_dE(orc, newp, mask):
    N, M = mask.shape
    out = np.zeros((N, M))
    orc.propose_with_dats(np.array([M], np.int64), np.zeros((N, 1), np.int64), newp.reshape(N, 3 * M),
                          mask.astype(np.int64), out)
    return out


def _cpu_pick(dE, mask, u, kT, rule):
    w = np.exp(-dE.ravel() / kT)
    if rule == "metropolis":
        w = np.minimum(w, 1.0)
    w = np.where(mask.ravel(), w, 0.0)
    c = np.cumsum(w)
    k = int(np.searchsorted(c, u * c[-1], side="right"))
    return min(k, len(w) - 1)


def _make(bc, seed, wrap):
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.bookkeeping import example_offsets
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    pos, q, E, rng = _dense_system(seed)
    offs = example_offsets(3.3)
    mm = 1.05 * np.max(np.abs(offs))
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=3, l=12, boundary_condition=bc, max_move=mm,
                 init_method="direct")
    kmc.initialise()
    lr = lr_operator_reference_layout(12, E) if bc == "pbc" else None
    orc = O.Oracle(pos, q, E, 3, 12, bc=bc, max_move=mm, lr=lr)
    orc.energy = orc.initialise()
    return kmc, orc, pos, q, E, offs, rng


@pytest.mark.parametrize("bc,wrap", (("pbc", True), ("free_space", False)))
def test_first_step_masks_and_dE(bc, wrap):
    from coulomb_kmc_amd.bookkeeping import LatticeKMC
    kmc, orc, pos, q, E, offs, rng = _make(bc, 31, wrap)
    u = rng.uniform(size=8)
    lk = LatticeKMC(kmc, offs, u, tol=0.01, wrap=wrap)
    lk.step(1)
    newp, ok = _cpu_proposals(pos, offs, E, wrap)
    mask = _cpu_masks(pos, newp, ok, E, 0.01, wrap)
    gm = lk.d_masks.cpu().numpy().reshape(mask.shape) > 0
    _print("masked", bc, int((~mask).sum()), "of", mask.size)
    assert (~mask).sum() > 0
    assert np.array_equal(gm, mask)
    assert np.array_equal(lk.d_new_pos.cpu().numpy().reshape(newp.shape), newp)
    dE = _cpu_dE(orc, newp, mask)
    g = lk.d_dE.cpu().numpy().reshape(mask.shape)
    err = rel_err(g[mask], dE[mask])
    _print("lattice dE", bc, err.max())
    assert err.max() <= 1e-10
    k = _cpu_pick(dE, mask, u[0], 1.0, "exp")
    assert lk.selected()[0] == k
    kmc.free()


@pytest.mark.parametrize("rule", ("exp", "metropolis"))
def test_step_sequence_matches_cpu(rule):
    from coulomb_kmc_amd.bookkeeping import LatticeKMC
    kmc, orc, pos, q, E, offs, rng = _make("pbc", 32, True)
    K = 12
    u = rng.uniform(size=K)
    lk = LatticeKMC(kmc, offs, u, tol=0.01, kT=1.0, rate=rule)
    lk.capture(steps_per_graph=5)       # one eager step, then 5 recorded
    lk.step(1)                          # eager again: 2 steps done
    lk.replay(2)                        # 12 steps
    assert lk.steps_done() == K
    sel = lk.selected()
    assert lk.sync_host() == 0
    M = len(offs)
    for step in range(K):
        cur = orc.positions()
        newp, ok = _cpu_proposals(cur, offs, E, True)
        mask = _cpu_masks(cur, newp, ok, E, 0.01, True)
        dE = _cpu_dE(orc, newp, mask)
        k = _cpu_pick(dE, mask, u[step], 1.0, rule)
        assert sel[step] == k, (step, sel[step], k)
        orc.accept(k // M, newp[k // M, k % M])
    gp, _ = kmc.get_device_particles()
    assert np.array_equal(gp, orc.positions())
    _print("lattice run", rule, kmc.energy, orc.energy)
    assert abs(kmc.energy - orc.energy) <= 1e-9 * abs(orc.energy)
    kmc.free()


def test_example_lattice_runs():
    """The example's own system: Ns^3 sites, all occupied, M = 14 offsets."""
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.bookkeeping import LatticeKMC, example_lattice, example_offsets
    pos, q, E = example_lattice(8, 3.3)
    offs = example_offsets(3.3)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=3, l=12, boundary_condition="pbc",
                 max_move=1.05 * np.max(np.abs(offs)))
    kmc.initialise()
    lk = LatticeKMC(kmc, offs, np.random.default_rng(5).uniform(size=20))
    lk.step(20)
    assert lk.steps_done() == 20
    assert lk.sync_host() == 0
    assert np.all(lk.selected() >= 0)
    e_dev = kmc.energy
    kmc2 = KMCFMM(kmc._pos().copy(), q.reshape(-1, 1), E, r=3, l=12, boundary_condition="pbc",
                  max_move=1.05 * np.max(np.abs(offs)))
    kmc2.initialise()
    _print("example lattice", e_dev, kmc2.energy)
    assert abs(e_dev - kmc2.energy) <= 1e-6 * abs(kmc2.energy)
    kmc.free()
    kmc2.free()
