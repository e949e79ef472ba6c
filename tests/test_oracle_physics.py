"""
Pins the oracle's propose/accept path on the CPU against independent physics and
the reference's own invariants:

* energies and dE against brute-force Coulomb sums (free space, '27') and an
  independent tinfoil Ewald sum (pbc) at the reference's tolerances
  (tests/host/test_free_space_1.py:150,230: 1e-5 at l=12; test_common.py:141-153: 1e-4);
* accept == re-initialise of the moved system (tests/host/test_free_space_1.py:450-542);
* cell tables == a Python model of swap-remove / append (coulomb_kmc/kmc_local.py:180-370);
* propose_with_dats == propose (tests/host/test_propose_dat_setup.py:447-);
* split identity pbc == '27' + 'ff-only' (tests/host/test_split_contrib.py:124-272).
"""

import numpy as np
import pytest

import oracle as O
from ewald import ewald_energy, free_space_energy, nearest27_energy
from helpers import neutral_system, ball_moves

from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout


def _system(seed, N=160, E=4.0):
    rng = np.random.RandomState(seed)
    pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
    q = np.array([(-1.0) ** (i + 1) for i in range(N)])
    q -= q.mean()
    return pos, q, rng


def _oracle(pos, q, E, bc, L=12, r=3, max_move=1.0):
    lr = lr_operator_reference_layout(L, E) if bc in ("pbc", "ff-only") else None
    o = O.Oracle(pos, q, E, r, L, bc=bc, max_move=max_move, lr=lr)
    o.initialise()
    return o


REFS = {
    "free_space": lambda p, q, E: free_space_energy(p, q),
    "27": lambda p, q, E: nearest27_energy(p, q, E),
    "pbc": lambda p, q, E: ewald_energy(p, q, E),
}
TOL = {"free_space": 1e-5, "27": 1e-5, "pbc": 1e-4}


@pytest.mark.parametrize("bc", ("free_space", "27", "pbc"))
def test_energy_and_dE_against_direct_sums(bc):
    E = 4.0
    pos, q, rng = _system(1234, E=E)
    o = _oracle(pos, q, E, bc)
    e_ref = REFS[bc](pos, q, E)
    assert abs(o.energy - e_ref) <= TOL[bc] * abs(e_ref)
    for _ in range(4):
        pid = rng.randint(0, len(q))
        mv = ball_moves(rng, pos, [pid], 2, 1.0, E, bc)[0][1]
        vals = o.propose(((pid, mv),))[0]
        for j in range(2):
            p2 = pos.copy()
            p2[pid] = mv[j]
            ref = REFS[bc](p2, q, E)
            # dE at the reference's tolerance on total energies
            assert abs(vals[j] - ref) <= TOL[bc] * abs(ref) + 2 * TOL[bc] * abs(e_ref)


def test_pbc_compact_cluster_tight():
    """Charges in a small central cluster: multipole truncation is negligible, so the
    full pbc energy and dE match the Ewald sum to 1e-9."""
    E = 6.0
    rng = np.random.RandomState(8657)
    N = 24
    pos = rng.uniform(-0.12 * E, 0.12 * E, (N, 3))
    q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    o = _oracle(pos, q, E, "pbc", L=20, r=3, max_move=0.5)
    e = ewald_energy(pos, q, E)
    assert abs(o.energy - e) <= 1e-9 * abs(e)
    pid = 5
    newp = pos[pid] + np.array([0.2, -0.1, 0.15])
    val = o.propose(((pid, newp[None, :]),))[0][0]
    p2 = pos.copy()
    p2[pid] = newp
    assert abs(val - ewald_energy(p2, q, E)) <= 1e-9 * abs(e)


@pytest.mark.parametrize("bc", ("free_space", "pbc"))
def test_accept_equals_reinitialise(bc):
    E = 4.0
    pos, q, rng = _system(8657, E=E)
    o = _oracle(pos, q, E, bc)
    pid = 11
    newp = ball_moves(rng, pos, [pid], 1, 1.0, E, bc)[0][1][0]
    o.accept(pid, newp)
    pos2 = pos.copy()
    pos2[pid] = newp
    o2 = _oracle(pos2, q, E, bc)
    a, b = o.get_store_expansions(), o2.get_store_expansions()
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b)) * 10
    if bc == "pbc":   # the level-0 multipole is only kept for the periodic far field
        M1, E1 = o.get_ME()
        M2, E2 = o2.get_ME()
        assert np.max(np.abs(M1 - M2)) <= 1e-13 * np.max(np.abs(M2))
    # energy accumulated from dE vs the re-initialised state: truncation-level agreement
    assert abs(o.energy - o2.energy) <= 1e-5 * abs(o2.energy)


def test_cell_tables_follow_swap_remove_append():
    E, bc = 4.0, "pbc"
    pos, q, rng = _system(99, E=E)
    o = _oracle(pos, q, E, bc)
    ncps = o.ncps
    w = E / ncps
    cell = lambda p: tuple(np.clip(((p + 0.5 * E) / w).astype(int), 0, ncps - 1))
    lin = lambda c: c[0] + ncps * (c[1] + ncps * c[2])
    model = {}
    for i in range(len(q)):
        model.setdefault(lin(cell(pos[i])), []).append(i)
    cur = pos.copy()
    for _ in range(25):
        pid = rng.randint(0, len(q))
        newp = ball_moves(rng, cur, [pid], 1, 1.0, E, bc)[0][1][0]
        oc, nc = lin(cell(cur[pid])), lin(cell(newp))
        lst = model[oc]
        k = lst.index(pid)
        lst[k] = lst[-1]
        lst.pop()
        model.setdefault(nc, []).append(pid)
        o.accept(pid, newp)
        cur[pid] = newp
    occ, ids, data = o.get_store()
    for c in range(ncps ** 3):
        assert occ[c] == len(model.get(c, []))
        assert list(ids[c, :occ[c]]) == model.get(c, [])
        for s in range(occ[c]):
            assert np.array_equal(data[c, s, :3], cur[ids[c, s]])


def test_propose_with_dats_equals_propose():
    E, bc = 4.0, "pbc"
    pos, q, rng = _system(4321, E=E)
    o = _oracle(pos, q, E, bc)
    N, M = len(q), 4
    site_max = np.array([1, 3], dtype=np.int64)
    sites = rng.randint(0, 2, size=(N, 1)).astype(np.int64)
    masks = (rng.uniform(size=(N, M)) < 0.25).astype(np.int64)
    ppos = np.zeros((N, 3 * M))
    for i in range(N):
        ppos[i] = ball_moves(rng, pos, [i], M, 1.0, E, bc)[0][1].reshape(-1)
    d = np.zeros((N, M))
    o.propose_with_dats(site_max, sites, ppos, masks, d)
    moves = []
    for i in range(N):
        taken = np.nonzero(masks[i])[0][:site_max[sites[i, 0]]]
        if len(taken):
            moves.append((i, ppos[i].reshape(M, 3)[taken]))
    vals = o.propose(tuple(moves))
    for (i, mv), v in zip(moves, vals):
        taken = np.nonzero(masks[i])[0][:site_max[sites[i, 0]]]
        assert np.max(np.abs((v - o.energy) - d[i, taken])) < 1e-12
    # untouched slots stay zero
    touched = np.zeros((N, M), bool)
    for i, mv in moves:
        touched[i, np.nonzero(masks[i])[0][:site_max[sites[i, 0]]]] = True
    assert np.all(d[~touched] == 0.0)


def test_split_identity_pbc_equals_27_plus_ff_only():
    E = 4.0
    pos, q, rng = _system(777, E=E)
    o_pbc = _oracle(pos, q, E, "pbc")
    o_27 = _oracle(pos, q, E, "27")
    o_ff = _oracle(pos, q, E, "ff-only")
    assert abs(o_pbc.energy - (o_27.energy + o_ff.energy)) <= 1e-13 * abs(o_pbc.energy) * 10
    pids = rng.choice(len(q), 10, replace=False)
    moves = ball_moves(rng, pos, pids, 3, 1.0, E, "pbc")
    a = np.concatenate(o_pbc.propose(moves)) - o_pbc.energy
    b = np.concatenate(o_27.propose(moves)) - o_27.energy
    c = np.concatenate(o_ff.propose(moves)) - o_ff.energy
    assert np.max(np.abs(a - (b + c))) <= 1e-12 * max(1.0, np.max(np.abs(a)))


def test_bad_moves_are_rejected():
    E = 4.0
    pos, q, rng = _system(5, E=E)
    o = _oracle(pos, q, E, "pbc", max_move=0.5)
    with pytest.raises(RuntimeError):
        o.propose(((0, np.array([[3.0, 0.0, 0.0]])),))
    with pytest.raises(RuntimeError):
        o.propose(((0, pos[0:1] + np.array([[0.6, 0.0, 0.0]])),))
