"""
GPU parity of injection and extraction (coulomb_kmc/kmc_inject_extract.py:82-533)
through the C ABI via KMCFMM:

* per-charge energies, group energies, propose_extract and propose_inject against the
  oracle on the same state (1e-10 relative; group energies 1e-12);
* after extract / inject the device state equals a fresh initialise of the new charge
  set: leaf expansions and multipole within 1e-11 of their scale, cell tables exact,
  the energy within FMM truncation (1e-6 relative), proposals on the new state
  against a fresh oracle at 1e-10;
* the reference's own sequences (tests/host/test_extract.py:311-400,
  tests/host/test_inject.py:171-270) against direct sums;
* DiscoverInjectExtract as tests/host/test_inject_extract.py:49-110.
"""

import ctypes

import numpy as np
import pytest

import oracle as O
from ewald import ewald_energy, free_space_energy, nearest27_energy
from helpers import neutral_system, ball_moves, rel_err

pytestmark = pytest.mark.gpu

BCS = ("free_space", "27", "pbc", "ff-only")
REFS = {
    "free_space": lambda p, q, E: free_space_energy(p, q),
    "27": lambda p, q, E: nearest27_energy(p, q, E),
    "pbc": lambda p, q, E: ewald_energy(p, q, E),
}


def _print(*a):
    print(*a, flush=True)


def _kmc(pos, q, E, bc, r=4, L=12, max_move=1.0):
    from coulomb_kmc_amd import KMCFMM
    k = KMCFMM(np.array(pos, dtype=np.float64), np.array(q, dtype=np.float64).reshape(-1, 1), E, r=r, l=L,
               boundary_condition=bc, max_move=max_move, init_method="direct")
    k.initialise()
    return k


def _oracle(pos, q, E, bc, r=4, L=12, max_move=1.0):
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    lr = lr_operator_reference_layout(L, E) if bc in ("pbc", "ff-only") else None
    o = O.Oracle(pos, np.asarray(q).ravel(), E, r, L, bc=bc, max_move=max_move, lr=lr)
    o.energy = o.initialise()
    return o


def _system(bc, N=300, E=6.0, seed=21):
    rng = np.random.default_rng(seed)
    pos, q = neutral_system(rng, N, E)
    return pos, q, rng


def _check_store(kmc):
    """Every charge in exactly one slot of its leaf, with its position and charge."""
    occ, ids, p4 = kmc.get_store()
    pos, cells = kmc.get_device_particles()
    n = kmc.N
    slot_cell, slot_id, slot_p = [], [], []
    for c in np.nonzero(occ)[0]:
        for s in range(occ[c]):
            slot_cell.append(c)
            slot_id.append(ids[c, s])
            slot_p.append(p4[c, s])
    slot_id = np.array(slot_id)
    assert len(slot_id) == n and np.array_equal(np.sort(slot_id), np.arange(n))
    order = np.argsort(slot_id)
    assert np.array_equal(np.array(slot_cell)[order], cells)
    sp = np.array(slot_p)[order]
    assert np.array_equal(sp[:, :3], pos)
    assert np.array_equal(sp[:, 3], kmc._q()[:n])
    assert np.array_equal(pos, kmc._pos()[:n])
    c3 = kmc.leaf_of(kmc._pos()[:n])
    nc = kmc._ncps
    assert np.array_equal(cells, c3[:, 0] + nc * (c3[:, 1] + nc * c3[:, 2]))


def _compare_fresh(kmc, bc, E, r=4, L=12):
    """Device state of kmc against a fresh initialise of its current charges."""
    pos, q = kmc._pos()[:kmc.N].copy(), kmc._q()[:kmc.N].copy()
    fresh = _kmc(pos, q, E, bc, r=r, L=L)
    if bc != "ff-only":
        a, b = kmc.get_expansions(), fresh.get_expansions()
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))
    if bc in ("pbc", "ff-only"):
        a, b = kmc.get_multipole(), fresh.get_multipole()
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))
    # energy bookkeeping (E + dE) against the energy of the new set, at FMM truncation
    assert abs(kmc.energy - fresh.energy) <= 1e-6 * max(1.0, abs(fresh.energy)), (kmc.energy, fresh.energy)
    assert kmc.energy == kmc.device_energy()
    fresh.free()
    _check_store(kmc)
    return pos, q


@pytest.mark.parametrize("bc", BCS)
def test_energies_match_oracle(bc):
    E = 6.0
    pos, q, rng = _system(bc)
    kmc, orc = _kmc(pos, q, E, bc), _oracle(pos, q, E, bc)
    ids = rng.choice(len(q), 40, replace=False)
    g = kmc.get_energy(ids.reshape(-1, 1))
    c = orc.charge_energy(ids)
    _print("charge energies", bc, rel_err(g, c).max())
    assert rel_err(g, c).max() <= 1e-10
    # rows and self energy
    rows = ids.reshape(8, 5)
    g = kmc.propose_extract(rows)
    c = orc.propose_extract(rows)
    _print("propose_extract", bc, rel_err(g, c).max())
    assert rel_err(g, c).max() <= 1e-10
    g = kmc.propose_extract(rows, self_energy=False)
    c = -np.array([orc.charge_energy(r).sum() for r in rows])
    assert rel_err(g, c).max() <= 1e-10
    # groups alone
    lib = kmc._lib
    import torch
    gp = rng.uniform(-0.5 * E, 0.5 * E, (6 * 4, 3))
    gq = np.tile([1.0, -1.0, 0.5, -0.5], 6)
    d_out = torch.zeros(6, dtype=torch.float64, device="cuda")
    tp, tq = torch.as_tensor(gp).cuda(), torch.as_tensor(gq).cuda()
    assert lib.kmc_group_energy(kmc.handle, kmc._stream(), 6, 4, tp.data_ptr(), tq.data_ptr(), d_out.data_ptr()) == 0
    g = d_out.cpu().numpy()
    c = np.array([orc.group_energy(gp[4 * i:4 * i + 4], gq[4 * i:4 * i + 4]) for i in range(6)])
    _print("group energies", bc, rel_err(g, c).max())
    assert rel_err(g, c).max() <= 1e-12
    assert abs(kmc.compute_energy(gp[:4], gq[:4]) - c[0]) <= 1e-12 * max(1.0, abs(c[0]))
    # injections
    ip = rng.uniform(-0.5 * E, 0.5 * E, (7, 3))
    iq = np.array([1.0, -1.0, 1.0, -1.0, 2.0, -1.0, -1.0])
    g = kmc.propose_inject(ip, iq)
    c = orc.propose_inject(ip, iq)
    _print("propose_inject", bc, g, c)
    assert abs(g - c) <= 1e-10 * max(1.0, abs(c))
    kmc.free()


@pytest.mark.parametrize("bc", BCS)
def test_extract_equals_fresh_initialise(bc):
    E = 6.0
    pos, q, rng = _system(bc, seed=22)
    kmc = _kmc(pos, q, E, bc)
    e0 = kmc.energy
    ids = rng.choice(len(q), 9, replace=False)
    d = kmc.propose_extract(ids)[0]
    kmc.extract(ids)
    assert kmc.N == len(q) - 9
    assert abs(kmc.energy - (e0 + d)) <= 1e-12 * max(1.0, abs(e0))
    # the same compaction on the host: descending ids, last charge into the hole
    p2, q2 = pos.copy(), q.copy()
    n = len(q)
    for i in sorted(ids.tolist(), reverse=True):
        p2[i], q2[i] = p2[n - 1], q2[n - 1]
        n -= 1
    assert np.array_equal(kmc._pos(), p2[:n]) and np.array_equal(kmc._q(), q2[:n])
    npos, nq = _compare_fresh(kmc, bc, E)
    # proposals on the new state against a fresh oracle
    orc = _oracle(npos, nq, E, bc)
    pids = rng.choice(kmc.N, 30, replace=False)
    moves = ball_moves(rng, npos, pids, 3, 1.0, E, bc)
    g = np.concatenate(kmc.propose(moves)) - kmc.energy
    c = np.concatenate(orc.propose(moves)) - orc.energy
    _print("propose after extract", bc, rel_err(g, c).max())
    assert rel_err(g, c).max() <= 1e-10
    kmc.free()


@pytest.mark.parametrize("bc", BCS)
def test_inject_equals_fresh_initialise(bc):
    E = 6.0
    pos, q, rng = _system(bc, seed=23)
    kmc = _kmc(pos, q, E, bc)
    e0 = kmc.energy
    # enough charges in one leaf to outgrow its slots, plus a few elsewhere
    w = E / 8
    centre = -0.5 * E + 2.5 * w
    cluster = centre + rng.uniform(-0.4 * w, 0.4 * w, (24, 3))
    other = rng.uniform(-0.5 * E, 0.5 * E, (5, 3))
    ip = np.concatenate([cluster, other])
    iq = np.where(np.arange(len(ip)) % 2 == 0, 1.0, -1.0)
    d = kmc.propose_inject(ip, iq)
    occ0 = kmc.get_store()[1].shape[1]
    kmc.inject({"positions": ip, "charges": iq.reshape(-1, 1)})
    assert kmc.get_store()[1].shape[1] > occ0
    assert kmc.N == len(q) + len(ip)
    assert abs(kmc.energy - (e0 + d)) <= 1e-12 * max(1.0, abs(e0))
    npos, nq = _compare_fresh(kmc, bc, E)
    orc = _oracle(npos, nq, E, bc)
    pids = np.concatenate([np.arange(len(q), kmc.N)[:10], rng.choice(len(q), 10, replace=False)])
    moves = ball_moves(rng, npos, pids, 3, 1.0, E, bc)
    g = np.concatenate(kmc.propose(moves)) - kmc.energy
    c = np.concatenate(orc.propose(moves)) - orc.energy
    _print("propose after inject", bc, rel_err(g, c).max())
    assert rel_err(g, c).max() <= 1e-10
    # accept a move of an injected charge, then the state again
    pid = len(q) + 3
    newp = moves[0][1][0] if moves[0][0] == pid else np.clip(npos[pid] + 0.3, -0.5 * E, 0.5 * E)
    kmc.accept((pid, newp))
    _check_store(kmc)
    kmc.free()


@pytest.mark.parametrize("bc", ("free_space", "pbc"))
def test_extract_inject_round_trip(bc):
    E = 6.0
    pos, q, rng = _system(bc, seed=24)
    kmc = _kmc(pos, q, E, bc)
    x0 = kmc.get_expansions() if bc != "ff-only" else None
    ids = rng.choice(len(q), 6, replace=False)
    p_out, q_out = pos[ids].copy(), q[ids].copy()
    kmc.extract(ids)
    kmc.inject({"positions": p_out, "charges": q_out.reshape(-1, 1)})
    x1 = kmc.get_expansions()
    assert np.max(np.abs(x1 - x0)) <= 1e-11 * np.max(np.abs(x0))
    # same charges (in another order): the state energy of a fresh start
    _compare_fresh(kmc, bc, E)
    kmc.free()


@pytest.mark.parametrize("bc", ("free_space", "27", "pbc"))
def test_reference_extract_sequence(bc):
    # tests/host/test_extract.py:311-400 (N=20, extent 2*3.1416, RandomState(3251), r=3, l=12)
    N, E = 20, 2 * 3.1416
    rng = np.random.RandomState(3251)
    pi = rng.uniform(low=-0.5 * E, high=0.5 * E, size=(N, 3))
    qi = np.array([(-1.0) ** px for px in range(N)])
    kmc = _kmc(pi, qi, E, bc, r=3, L=12)
    direct = REFS[bc]
    assert abs(kmc.energy - direct(pi, qi, E)) / abs(direct(pi, qi, E)) < 1e-5
    for _ in range(9):
        n = kmc.N
        qv = kmc._q()
        inds, available = [], set(range(n))
        for sign in (1, -1):
            ind = rng.randint(0, n)
            while (sign * qv[ind] < 0) or (ind not in available):
                ind = rng.randint(0, n)
            inds.append(ind)
            available.remove(ind)
        phi0 = kmc.energy
        d = kmc.propose_extract(inds)[0]
        kmc.extract(inds)
        assert abs(kmc.energy - phi0 - d) < 1e-8
        correct = direct(kmc._pos().copy(), kmc._q().copy(), E)
        assert abs(correct - kmc.energy) < 1e-5, (correct, kmc.energy)
    kmc.free()


@pytest.mark.parametrize("bc", ("free_space", "27", "pbc"))
def test_reference_inject_sequence(bc):
    # tests/host/test_inject.py:171-270: a system without some charges, then inject them
    N, E = 20, 2 * 3.1416
    rng = np.random.RandomState(3251)
    pi = rng.uniform(low=-0.5 * E, high=0.5 * E, size=(N, 3))
    qi = np.array([(-1.0) ** px for px in range(N)])
    direct = REFS[bc]
    phi_full = direct(pi, qi, E)
    for _ in range(4):
        num = rng.randint(1, 5)
        inds, available = [], set(range(N))
        for sign in (1, -1):
            for _k in range(num):
                ind = rng.randint(0, N)
                while (sign * qi[ind] < 0) or (ind not in available):
                    ind = rng.randint(0, N)
                inds.append(ind)
                available.remove(ind)
        keep = np.array(sorted(available))
        kmc = _kmc(pi[keep], qi[keep], E, bc, r=3, L=12)
        kmc.inject({"positions": pi[inds], "charges": qi[inds].reshape(-1, 1)})
        assert abs(kmc.energy - phi_full) / abs(phi_full) < 1e-5, (kmc.energy, phi_full)
        kmc.free()


def test_reinitialise_with_another_charge_count():
    E = 6.0
    pos, q, rng = _system("pbc", seed=25)
    kmc = _kmc(pos[:200], q[:200], E, "pbc")
    kmc.positions = pos.copy()
    kmc.charges = q.reshape(-1, 1).copy()
    kmc.initialise()
    fresh = _kmc(pos, q, E, "pbc")
    assert kmc.N == len(q)
    assert abs(kmc.energy - fresh.energy) <= 1e-12 * abs(fresh.energy)
    kmc.free()
    fresh.free()


def test_inject_extract_errors():
    from coulomb_kmc_amd._native import KMCError
    E = 6.0
    pos, q, rng = _system("pbc", N=100, seed=26)
    kmc = _kmc(pos, q, E, "pbc")
    with pytest.raises(KMCError):
        kmc.extract([3, 3])
    with pytest.raises(KMCError):
        kmc.extract([100])
    with pytest.raises(KMCError):
        kmc.propose_extract([[0, 1], [2, 500]])
    with pytest.raises(RuntimeError):
        kmc.inject({"positions": np.array([[0.0, 0.0, 3.5]]), "charges": np.array([[1.0]])})
    assert kmc.N == 100
    _check_store(kmc)
    kmc.free()


def test_discover_sites():
    # tests/host/test_inject_extract.py:49-110
    from coulomb_kmc_amd import DiscoverInjectExtract
    E = 4.0
    rng = np.random.RandomState(19)
    N = 100
    pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
    s = rng.permutation(N)
    inject_sites = rng.uniform(-0.5 * E, 0.5 * E, (20, 3))
    inject_sites[:10] = pos[s[:10]]
    extract_sites = pos[s[10:20]].copy()
    flag = np.zeros((N, 1), np.int64)
    die = DiscoverInjectExtract(inject_sites, extract_sites, pos, flag)
    iflags = die()
    assert np.array_equal(iflags[:10], np.ones(10, np.int64))
    assert np.array_equal(iflags[10:], np.zeros(10, np.int64))
    for px in range(N):
        if px in s[10:20]:
            assert flag[px, 0] == 1 + int(np.nonzero(s[10:20] == px)[0][0])
        else:
            assert flag[px, 0] == 0
