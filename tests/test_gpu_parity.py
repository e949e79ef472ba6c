"""
GPU parity: the HIP path (through the C ABI via KMCFMM) against the oracle
(oracle/kmc_oracle.cpp, the CPU restatement of the reference) on the same seeded
inputs.  Bars: FP64 dE within 1e-10 relative, mixed within 1e-8 relative, cell
tables and accepted-move sequences bit-exact, expansions after accepts within
1e-11 of their scale.
"""

import numpy as np
import pytest

import oracle as O
from helpers import neutral_system, ball_moves, rel_err, metropolis_pick

pytestmark = pytest.mark.gpu

BCS = ("free_space", "27", "pbc", "ff-only")


def make_pair(bc, N=300, E=6.0, r=4, L=12, max_move=1.0, seed=11):
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    rng = np.random.default_rng(seed)
    pos, q = neutral_system(rng, N, E)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=max_move)
    kmc.initialise()
    lr = lr_operator_reference_layout(L, E) if bc in ("pbc", "ff-only") else None
    orc = O.Oracle(pos, q, E, r, L, bc=bc, max_move=max_move, lr=lr)
    orc.initialise()
    return kmc, orc, pos, q, rng


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("L", (8, 12))
def test_initialise_matches_oracle(bc, L):
    kmc, orc, pos, q, rng = make_pair(bc, L=L)
    _print("init", bc, L, kmc.energy, orc.energy)
    assert abs(kmc.energy - orc.energy) <= 1e-11 * max(1.0, abs(orc.energy))
    if bc != "ff-only":
        eg, ec = kmc.get_expansions(), orc.get_expansions()
        assert np.max(np.abs(eg - ec)) <= 1e-12 * np.max(np.abs(ec))
    if bc in ("pbc", "ff-only"):
        Mg = kmc.get_multipole()
        Mc, Ec = orc.get_ME()
        assert np.max(np.abs(Mg - Mc)) <= 1e-12 * np.max(np.abs(Mc))
        assert np.max(np.abs(Mg - Ec)) <= 1e-12 * np.max(np.abs(Ec))
    kmc.free()


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("prec", ("fp64", "mixed"))
def test_propose_matches_oracle(bc, prec):
    kmc, orc, pos, q, rng = make_pair(bc)
    kmc.precision = prec
    pids = rng.choice(len(q), 40, replace=False)
    moves = ball_moves(rng, pos, pids, 4, 1.0, 6.0, bc)
    g = np.concatenate(kmc.propose(moves)) - kmc.energy
    c = np.concatenate(orc.propose(moves)) - orc.energy
    err = rel_err(g, c)
    _print("propose", bc, prec, "max rel", err.max(), "median", np.median(err))
    assert err.max() <= (1e-10 if prec == "fp64" else 1e-8)
    kmc.free()


@pytest.mark.parametrize("L", (4, 10, 16, 20))
def test_propose_orders(L):
    kmc, orc, pos, q, rng = make_pair("pbc", N=200, E=5.0, r=3, L=L)
    pids = rng.choice(len(q), 20, replace=False)
    moves = ball_moves(rng, pos, pids, 3, 1.0, 5.0, "pbc")
    g = np.concatenate(kmc.propose(moves)) - kmc.energy
    c = np.concatenate(orc.propose(moves)) - orc.energy
    err = rel_err(g, c)
    _print("orders", L, err.max())
    assert err.max() <= 1e-10
    kmc.free()


@pytest.mark.parametrize("bc", ("free_space", "pbc"))
def test_propose_with_dats_matches_oracle(bc):
    kmc, orc, pos, q, rng = make_pair(bc)
    N, M = len(q), 6
    site_max = np.array([2, 4, 6], dtype=np.int64)
    sites = rng.integers(0, 3, size=(N, 1)).astype(np.int64)
    masks = (rng.uniform(size=(N, M)) < 0.3).astype(np.int64)
    ppos = np.zeros((N, 3 * M))
    for i in range(N):
        mv = ball_moves(rng, pos, [i], M, 1.0, 6.0, bc)[0][1]
        ppos[i] = mv.reshape(-1)
    dg = np.full((N, M), -7.0)
    dc = np.full((N, M), -7.0)
    kmc.propose_with_dats(site_max, sites, ppos, masks, dg)
    orc.propose_with_dats(site_max, sites, ppos, masks, dc)
    touched = dc != -7.0
    assert np.array_equal(touched, dg != -7.0)
    err = rel_err(dg[touched], dc[touched])
    _print("dats", bc, touched.sum(), err.max())
    assert err.max() <= 1e-10
    # propose_with_dats == propose on the same moves (tests/host/test_propose_dat_setup.py:447-)
    moves = []
    for i in range(N):
        taken = np.nonzero(masks[i])[0][: site_max[sites[i, 0]]]
        if len(taken):
            moves.append((i, ppos[i].reshape(M, 3)[taken]))
    e = kmc.propose(tuple(moves))
    for (i, mv), vals in zip(moves, e):
        taken = np.nonzero(masks[i])[0][: site_max[sites[i, 0]]]
        assert np.allclose(vals - kmc.energy, dg[i, taken], rtol=1e-12, atol=1e-12)
    kmc.free()


def _store_equal(kmc, orc):
    occ_g, ids_g, p4_g = kmc.get_store()
    occ_c, ids_c, data_c = orc.get_store()
    assert np.array_equal(occ_g.astype(np.int64), occ_c)
    for c in np.nonzero(occ_c)[0]:
        n = occ_c[c]
        assert np.array_equal(ids_g[c, :n].astype(np.int64), ids_c[c, :n])
        assert np.array_equal(p4_g[c, :n, :3], data_c[c, :n, :3])
        assert np.array_equal(p4_g[c, :n, 3], data_c[c, :n, 3])


@pytest.mark.parametrize("bc", BCS)
def test_accept_sequence_matches_oracle(bc):
    kmc, orc, pos, q, rng = make_pair(bc, N=240, E=6.0, r=4, L=10)
    picks_g, picks_c = [], []
    for step in range(12):
        pg = kmc._pos()
        pids = rng.choice(len(q), 16, replace=False)
        moves = ball_moves(rng, pg, pids, 4, 1.0, 6.0, bc)
        dg = np.concatenate(kmc.propose(moves)) - kmc.energy
        dc = np.concatenate(orc.propose(moves)) - orc.energy
        u = rng.uniform()
        kg, kc = metropolis_pick(dg, u), metropolis_pick(dc, u)
        picks_g.append(kg)
        picks_c.append(kc)
        flat = [(pid, p) for pid, mv in moves for p in mv]
        pid, npos = flat[kg]
        kmc.accept((pid, npos))
        orc.accept(pid, npos)
        assert abs(kmc.energy - orc.energy) <= 1e-10 * max(1.0, abs(orc.energy))
    assert picks_g == picks_c
    pos_g, cells_g = kmc.get_device_particles()
    assert np.array_equal(pos_g, orc.positions())
    assert np.array_equal(cells_g.astype(np.int64), orc.cells())
    assert np.array_equal(kmc._pos(), orc.positions())
    if bc != "ff-only":
        _store_equal(kmc, orc)
        eg, ec = kmc.get_expansions(), orc.get_expansions()
        err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
        _print("accept", bc, "expansion err", err)
        assert err <= 1e-11
    if bc in ("pbc", "ff-only"):
        Mg = kmc.get_multipole()
        Mc, _ = orc.get_ME()
        assert np.max(np.abs(Mg - Mc)) <= 1e-12 * np.max(np.abs(Mc))
    kmc.free()


@pytest.mark.parametrize("bc", ("pbc", "free_space"))
@pytest.mark.parametrize("L", (8, 11, 16))
def test_every_tile_shape_matches_oracle(bc, L):
    """Every accept-sweep tile shape (column pairs per lane, split or fused workgroup
    layout with the LDS-staged update, image split over waves) and every near-field
    lane count gives the oracle's dE and, after an accept, its expansions (one system,
    one propose + accept per shape, mirrored on the oracle)."""
    from coulomb_kmc_amd import tune
    shapes = tune.sweep_candidates(bc != "free_space")
    near = tune.NEAR_LANES
    kmc, orc, pos, q, rng = make_pair(bc, N=300, E=6.0, r=4, L=L, seed=100 + L)
    for i, (npair, block, split) in enumerate(shapes):
        kmc.set_tiles(near_lanes=near[i % len(near)], sweep_block=block, sweep_pairs=npair, sweep_split=split)
        pids = rng.choice(len(q), 4, replace=False)
        moves = ball_moves(rng, kmc._pos(), pids, 3, 1.0, 6.0, bc)
        dg = np.concatenate(kmc.propose(moves)) - kmc.energy
        dc = np.concatenate(orc.propose(moves)) - orc.energy
        assert np.max(np.abs(dg - dc) / np.maximum(np.abs(dc), 1e-3)) <= 1e-10, (npair, block, split)
        pid, npos = moves[0][0], moves[0][1][1]
        kmc.accept((pid, npos))
        orc.accept(pid, npos)
        eg, ec = kmc.get_expansions(), orc.get_expansions()
        err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
        _print("tiles", bc, L, (npair, block, split), near[i % len(near)], "expansion err", err)
        assert err <= 1e-11, (npair, block, split)
    kmc.free()


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("r", (2, 4))
def test_grouped_near_field(bc, r):
    """The grouped near-field kernel (near_lanes 0: a moving charge's old position and
    proposals over one box of leaves) against the per-point kernel on the same state
    (same pairs, other summation order) and against the oracle: 1 to 21 proposals per
    charge (two passes above 15), proposals up to 2.7 leaves away (the per-point
    fallback), two leaves per side (periodic images of the same leaf)."""
    kmc, orc, pos, q, rng = make_pair(bc, N=300, E=6.0, r=r, L=8, max_move=2.0, seed=40 + r)
    pids = rng.choice(len(q), 30, replace=False)
    moves = []
    for i, pid in enumerate(pids):
        moves += list(ball_moves(rng, pos, [pid], 1 + (i * 7) % 21, 2.0, 6.0, bc))
    moves = tuple(moves)
    for prec, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
        kmc.precision = prec
        kmc.set_tiles(near_lanes=27)
        ref = np.concatenate(kmc.propose(moves)) - kmc.energy
        kmc.set_tiles(near_lanes=0)
        g = np.concatenate(kmc.propose(moves)) - kmc.energy
        c = np.concatenate(orc.propose(moves)) - orc.energy
        _print("grouped near", bc, r, prec, rel_err(g, ref).max(), rel_err(g, c).max())
        assert rel_err(g, ref).max() <= 1e-11
        assert rel_err(g, c).max() <= tol
    kmc.free()


def test_step_counter_below_zero_wraps():
    """The synthetic move stream and the selection take row (counter mod nsteps) of
    their per-step tables, counted from 0 also for counters below zero (a counter of -1
    once addressed the row before the table)."""
    import torch
    kmc, orc, pos, q, rng = make_pair("pbc", N=200, E=5.0, r=3, L=8)
    nsteps, P, k = 5, 4, 2
    pids = np.stack([rng.choice(200, P, replace=False) for _ in range(nsteps)]).astype(np.int64)
    disp = rng.uniform(-0.3, 0.3, (nsteps, P * k, 3))
    t_pids = torch.as_tensor(pids).cuda()
    t_disp = torch.as_tensor(disp).cuda()
    d_pid = torch.zeros(P, dtype=torch.int64, device="cuda")
    d_es = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
    d_np = torch.zeros(P * k * 3, dtype=torch.float64, device="cuda")
    for c in (-1, -7, 12):
        counter = torch.full((1,), c, dtype=torch.int64, device="cuda")
        kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
        assert np.array_equal(d_pid.cpu().numpy(), pids[c % nsteps]), c
    assert kmc.check_flags() == 0
    kmc.free()


def test_accept_movedata_equals_accept():
    """kmc_accept_movedata (the reference's 10-word move record, kmc_fmm.py:811-836)
    does what kmc_accept does, and refuses a record that does not match the state."""
    import ctypes
    a, orc, pos, q, rng = make_pair("pbc", N=200, E=5.0, r=3, L=8, seed=31)
    b, _, _, _, _ = make_pair("pbc", N=200, E=5.0, r=3, L=8, seed=31)
    n = a._ncps
    pid = 11
    newp = np.clip(pos[pid] + np.array([0.4, -0.3, 0.2]), -2.5, 2.5)
    c3o, c3n = a.leaf_of(pos[pid:pid + 1])[0], a.leaf_of(newp[None, :])[0]
    md = np.zeros(10, np.int64)
    rd = md[:7].view(np.float64)
    rd[0:3], rd[3:6], rd[6] = pos[pid], newp, q[pid]
    md[7] = pid
    md[8] = c3o[0] + n * (c3o[1] + n * c3o[2])
    md[9] = c3n[0] + n * (c3n[1] + n * c3n[2])
    e = ctypes.c_double(0.0)
    bad = md.copy()
    bad[9] = (md[9] + 1) % n ** 3
    assert a._lib.kmc_accept_movedata(a.handle, a._stream(), bad.ctypes.data_as(ctypes.c_void_p), ctypes.byref(e)) != 0
    assert a._lib.kmc_accept_movedata(a.handle, a._stream(), md.ctypes.data_as(ctypes.c_void_p), ctypes.byref(e)) == 0
    b.accept((pid, newp))
    assert e.value == b.energy
    assert np.array_equal(a.get_expansions(), b.get_expansions())
    assert np.array_equal(a.get_device_particles()[0], b.get_device_particles()[0])
    a.free()
    b.free()


def test_accept_equals_reinitialise():
    """Sweep after a move == fresh initialise of the moved system (direct P2L, exact)."""
    from coulomb_kmc_amd import KMCFMM
    kmc, orc, pos, q, rng = make_pair("pbc", N=200, E=5.0, r=4, L=12)
    pid = 17
    newp = kmc._pos()[pid] + np.array([0.3, -0.2, 0.25])
    kmc.accept((pid, newp))
    pos2 = pos.copy()
    pos2[pid] = newp
    k2 = KMCFMM(pos2.copy(), q.reshape(-1, 1), 5.0, r=4, l=12, boundary_condition="pbc", max_move=1.0)
    k2.initialise()
    a, b = kmc.get_expansions(), k2.get_expansions()
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    # accumulated energy vs fresh state energy: equal up to expansion truncation
    assert abs(kmc.energy - k2.energy) <= 1e-4 * abs(k2.energy)
    kmc.free()
    k2.free()


def test_device_step_matches_host_path():
    """kmc_step (select + accept on the device) == host propose/select/accept."""
    import torch
    from coulomb_kmc_amd import KMCFMM
    bc, N, E, r, L = "pbc", 300, 6.0, 4, 12
    rng = np.random.default_rng(5)
    pos, q = neutral_system(rng, N, E)
    a = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=1.0)
    b = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=1.0)
    a.initialise()
    b.initialise()
    steps, P, k = 10, 24, 4
    T = P * k
    pids_all = np.stack([rng.choice(N, P, replace=False) for _ in range(steps)]).astype(np.int64)
    d = rng.normal(size=(steps, T, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    d *= rng.uniform(size=(steps, T, 1)) ** (1.0 / 3.0)
    us = rng.uniform(size=steps)
    dev = lambda x, t: torch.as_tensor(x, dtype=t).cuda()
    t_pids, t_disp, t_u = dev(pids_all, torch.int64), dev(d, torch.float64), dev(us, torch.float64)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    sel = torch.full((steps,), -1, dtype=torch.int64, device="cuda")
    d_pid = torch.zeros(P, dtype=torch.int64, device="cuda")
    d_es = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
    d_np = torch.zeros(T * 3, dtype=torch.float64, device="cuda")
    d_dE = torch.zeros(T, dtype=torch.float64, device="cuda")
    a.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
    a.propose_device(d_pid, d_es, d_np, d_dE)    # sizes the work buffers
    for s in range(steps):
        a.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
        a.step_device(d_pid, d_es, d_np, d_dE, t_u, 1.0, counter, sel)
        # host path on b
        pb = b._pos()
        moves = []
        for j in range(P):
            pid = int(pids_all[s, j])
            npos = pb[pid] + d[s, j * k:(j + 1) * k]
            npos = np.where(npos < -0.5 * E, npos + E, npos)
            npos = np.where(npos >= 0.5 * E, npos - E, npos)
            moves.append((pid, npos))
        db = np.concatenate(b.propose(tuple(moves))) - b.energy
        kb = metropolis_pick(db, us[s])
        flat = [(pid, p) for pid, mv in moves for p in mv]
        b.accept(flat[kb])
        torch.cuda.synchronize()
        assert int(sel[s].item()) == kb, (s, int(sel[s].item()), kb)
    assert a.check_flags() == 0
    assert abs(a.device_energy() - b.energy) <= 1e-10 * abs(b.energy)
    pa, ca = a.get_device_particles()
    pb, cb = b.get_device_particles()
    assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
    assert np.max(np.abs(a.get_expansions() - b.get_expansions())) == 0.0
    a.free()
    b.free()


def test_free_space_against_direct_sum():
    """dE vs brute-force Coulomb differences at the reference's own tolerance
    (tests/host/test_free_space_1.py:194-230: 1e-5 at l=12)."""
    from ewald import free_space_energy
    kmc, orc, pos, q, rng = make_pair("free_space", N=200, E=4.0, r=3, L=12, max_move=4.0)
    e0 = free_space_energy(pos, q)
    assert abs(kmc.energy - e0) <= 1e-5 * abs(e0)
    for _ in range(5):
        pid = int(rng.integers(0, len(q)))
        newp = rng.uniform(-2.0, 2.0, (3, 3))
        vals = kmc.propose(((pid, newp),))[0]
        for j in range(3):
            p2 = pos.copy()
            p2[pid] = newp[j]
            ref = free_space_energy(p2, q)
            assert abs(vals[j] - ref) <= 1e-5 * abs(ref)
    kmc.free()


def test_bad_move_raises():
    kmc, orc, pos, q, rng = make_pair("pbc")
    with pytest.raises(RuntimeError):
        kmc.propose(((0, np.array([[10.0, 0.0, 0.0]])),))
    with pytest.raises(RuntimeError):
        kmc.propose(((0, (pos[0] + np.array([1.5, 0, 0]))[None, :]),))
    kmc.free()


def test_device_calls_before_particles_raise():
    """Entry points that touch the leaf expansions or the particle arrays refuse to run
    before the particles are set (the host check returns before any launch)."""
    import torch
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(0)
    pos, q = neutral_system(rng, 50, 4.0)
    kmc = KMCFMM(pos, q, 4.0, r=3, l=8, boundary_condition="pbc")
    with pytest.raises(RuntimeError):
        kmc.sweep_hop(1.0, pos[0], pos[1])
    with pytest.raises(RuntimeError):
        kmc.get_expansions()
    z = torch.zeros(8, dtype=torch.int64, device="cuda")
    d = torch.zeros(24, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):
        kmc.materialize_moves(1, 1, z, d, z, z, z, d)
    kmc.free()


def test_use_before_initialise_raises():
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(0)
    pos, q = neutral_system(rng, 50, 4.0)
    kmc = KMCFMM(pos, q, 4.0, r=3, l=8, boundary_condition="free_space")
    with pytest.raises(RuntimeError):
        kmc.propose(((0, pos[0:1]),))
    kmc.free()


@pytest.mark.parametrize("bc,L", [("pbc", 16), ("free_space", 16), ("pbc", 15)])
def test_blocked_sweep_truncation_matches_oracle(bc, L):
    """The blocked accept sweep (sweep_block 1: a 4x4x4 block of leaves per workgroup,
    orders of each source image dropped where they change the block's potential by less
    than 2^-60 per unit charge) in a box wide enough that most images lose orders:
    expansions after accepts against the oracle's full P2L sweep (1e-11 of their
    scale), then dE of new proposals against the extended reference (1e-10).  L = 15
    and 16 take the truncated buckets' images in wave pairs (each wave one source, all
    four column pairs of its group; L = 15 with a one-column middle pair)."""
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    N, E, r = 3000, 60.0, 5
    rng = np.random.default_rng(77)
    pos, q = neutral_system(rng, N, E)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=1.5,
                 init_method="direct")
    kmc.set_tiles(near_lanes=4, sweep_block=1, sweep_pairs=2, sweep_split=1)
    kmc.initialise()
    lr = lr_operator_reference_layout(L, E) if bc == "pbc" else None
    orc = O.Oracle(pos, q, E, r, L, bc=bc, max_move=1.5, lr=lr)
    orc.init_from_expansions(kmc.get_expansions(), kmc.energy)
    for step in range(6):
        pid = int(rng.integers(N))
        newp = ball_moves(rng, kmc._pos(), [pid], 1, 1.5, E, bc)[0][1][0]
        kmc.accept((pid, newp))
        orc.accept(pid, newp)
    eg, ec = kmc.get_expansions(), orc.get_expansions()
    err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
    _print("blocked sweep", bc, "expansion err after 6 accepts", err)
    assert err <= 1e-11
    pids = rng.choice(N, 64, replace=False)
    moves = ball_moves(rng, kmc._pos(), pids, 4, 1.5, E, bc)
    g = np.concatenate(kmc.propose(moves)) - kmc.energy
    es = np.arange(65, dtype=np.int64) * 4
    ext = orc.propose_check(pids, es, np.concatenate([m[1] for m in moves]), extended=True)
    _print("blocked sweep", bc, "dE vs extended", rel_err(g, ext).max())
    assert rel_err(g, ext).max() <= O.DE_BAR_FP64
    kmc.free()


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("r,N,E", [(2, 300, 6.0), (3, 300, 6.0), (4, 600, 8.0), (3, 1500, 6.0)])
def test_tiled_near_field(bc, r, N, E):
    """The leaf-tiled near field (near_lanes -G: points counting-sorted by 4x4x4 leaf
    tile, one workgroup per tile with the tile's 6^3 leaves staged in LDS) against the
    per-point kernel with the same lane count and against the oracle: a 2-leaf grid (not
    tiled: the per-point kernel runs), a 4-leaf grid (one tile whose halo wraps onto
    itself), an 8-leaf grid, and 1500 charges on 4^3 leaves (more records than the LDS
    holds: the tile is evaluated from memory)."""
    kmc, orc, pos, q, rng = make_pair(bc, N=N, E=E, r=r, L=8, max_move=1.0, seed=60 + r)
    pids = rng.choice(len(q), 40, replace=False)
    moves = ball_moves(rng, pos, pids, 5, 1.0, E, bc)
    for G in (4, 8):
        for prec, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
            kmc.precision = prec
            kmc.set_tiles(near_lanes=G)
            ref = np.concatenate(kmc.propose(moves)) - kmc.energy
            kmc.set_tiles(near_lanes=-G)
            g = np.concatenate(kmc.propose(moves)) - kmc.energy
            c = np.concatenate(orc.propose(moves)) - orc.energy
            _print("tiled near", bc, r, N, G, prec, rel_err(g, ref).max(), rel_err(g, c).max())
            assert rel_err(g, ref).max() <= 1e-11
            assert rel_err(g, c).max() <= tol
    kmc.free()


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("r,N,E", [(2, 300, 6.0), (3, 300, 6.0), (5, 3000, 16.0)])
def test_sorted_near_field(bc, r, N, E):
    """The per-point near field over the points sorted by leaf tile (near_order 1, the
    blocks remapped so that each XCD takes one contiguous range) equals the per-point
    kernel in the proposals' order bit for bit, for 1, 4, 8 and 16 lanes per point (1 lane:
    fewer than 8 workgroups; 16 lanes: a block count that is not a multiple of 8), on a
    2-leaf grid (not tiled: order 0 runs), one tile and 8^3 tiles; and the oracle's
    extended-precision checker to the parity bars."""
    kmc, orc, pos, q, rng = make_pair(bc, N=N, E=E, r=r, L=8, max_move=1.0, seed=70 + r)
    pids = rng.choice(len(q), min(200, len(q) // 2), replace=False)
    moves = ball_moves(rng, pos, pids, 5, 1.0, E, bc)
    # the extended-precision checker on the GPU's own leaf expansions (the oracle's
    # independently initialised ones differ by their FP64 rounding, ~1e-13 absolute,
    # which the relative measure's 1e-3 floor would count)
    orc.init_from_expansions(kmc.get_expansions(), kmc.energy)
    c = orc.propose_check(pids, np.arange(len(pids) + 1, dtype=np.int64) * 5, np.concatenate([m[1] for m in moves]),
                          extended=True)
    for G in (1, 4, 8, 16):
        for prec, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
            kmc.precision = prec
            kmc.set_tiles(near_lanes=G, near_order=0)
            ref = np.concatenate(kmc.propose(moves))
            kmc.set_tiles(near_lanes=G, near_order=1)
            g = np.concatenate(kmc.propose(moves))
            _print("sorted near", bc, r, N, G, prec, rel_err(g - kmc.energy, c).max())
            assert np.array_equal(g, ref)
            assert rel_err(g - kmc.energy, c).max() <= tol
    kmc.free()


def test_propose_with_dats_many_scan_blocks_and_bad_site():
    """propose_with_dats over 5000 charges (the mask compaction's scan spans several
    workgroups) equals propose on the same moves; a current site outside
    site_max_counts is refused on the device (KMC_E_ARG) through the C ABI."""
    import ctypes
    import torch
    from coulomb_kmc_amd import _native
    kmc, orc, pos, q, rng = make_pair("pbc", N=5000, E=14.0, r=4, L=8, max_move=1.0, seed=91)
    N, M = len(q), 4
    site_max = np.array([1, 2, 4], dtype=np.int64)
    sites = rng.integers(0, 3, size=(N, 1)).astype(np.int64)
    masks = (rng.uniform(size=(N, M)) < 0.5).astype(np.int64)
    ppos = np.zeros((N, 3 * M))
    for i in range(N):
        ppos[i] = ball_moves(rng, pos, [i], M, 1.0, 14.0, "pbc")[0][1].reshape(-1)
    dg = np.full((N, M), -7.0)
    kmc.propose_with_dats(site_max, sites, ppos, masks, dg)
    moves = []
    for i in range(N):
        taken = np.nonzero(masks[i])[0][: site_max[sites[i, 0]]]
        if len(taken):
            moves.append((i, ppos[i].reshape(M, 3)[taken]))
    e = kmc.propose(tuple(moves))
    for (i, mv), vals in zip(moves, e):
        taken = np.nonzero(masks[i])[0][: site_max[sites[i, 0]]]
        assert np.allclose(vals - kmc.energy, dg[i, taken], rtol=1e-12, atol=1e-12)
    dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
    bad = sites.copy()
    bad[17, 0] = 3
    d_smc, d_sites = dev(site_max, torch.int64), dev(bad.ravel(), torch.int64)
    d_ppos, d_masks = dev(ppos, torch.float64), dev(masks, torch.int64)
    d_out = torch.zeros((N, M), dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _native.lib().kmc_propose_with_dats(kmc.handle, st, M, d_smc.data_ptr(), 3, d_sites.data_ptr(),
                                             d_ppos.data_ptr(), d_masks.data_ptr(), None, d_out.data_ptr(), 0, None)
    assert rc == _native.KMC_E_ARG
    # the error counter is back at zero: the next call works
    rc = _native.lib().kmc_propose_with_dats(kmc.handle, st, M, d_smc.data_ptr(), 3,
                                             dev(sites.ravel(), torch.int64).data_ptr(), d_ppos.data_ptr(),
                                             d_masks.data_ptr(), None, d_out.data_ptr(), 0, None)
    assert rc == 0
    kmc.free()


def test_device_step_refuses_bad_move():
    """A device step whose proposals include one beyond max_move accepts nothing and
    raises flag 32; the next step with good moves goes on as usual."""
    import torch
    kmc, orc, pos, q, rng = make_pair("pbc", N=300, E=6.0, r=4, L=8, max_move=0.5, seed=92)
    P, k = 8, 2
    pids = rng.choice(len(q), P, replace=False)
    good = ball_moves(rng, pos, pids, k, 0.4, 6.0, "pbc")
    npos = np.concatenate([m[1] for m in good])
    dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
    d_pid, d_es = dev(pids, torch.int64), dev(np.arange(P + 1) * k, torch.int64)
    d_dE = torch.zeros(P * k, dtype=torch.float64, device="cuda")
    us = dev(np.array([0.3, 0.6]), torch.float64)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    sel = torch.full((2,), -7, dtype=torch.int64, device="cuda")
    kmc.propose_device(d_pid, d_es, dev(npos, torch.float64), d_dE)   # sizes the buffers
    e0 = kmc.device_energy()
    p0, _ = kmc.get_device_particles()
    bad = npos.copy()
    bad[3] = pos[pids[1]] + np.array([0.9, 0.0, 0.0])                 # beyond max_move 0.5
    bad[3] = np.where(bad[3] >= 3.0, bad[3] - 6.0, bad[3])
    kmc.step_device(d_pid, d_es, dev(bad, torch.float64), d_dE, us, 1.0, counter, sel)
    torch.cuda.synchronize()
    assert kmc.check_flags() == 32
    assert int(sel[0].item()) == -1
    assert kmc.device_energy() == e0
    p1, _ = kmc.get_device_particles()
    assert np.array_equal(p0, p1)
    kmc.step_device(d_pid, d_es, dev(npos, torch.float64), d_dE, us, 1.0, counter, sel)
    torch.cuda.synchronize()
    assert kmc.check_flags() == 0
    assert int(sel[1].item()) >= 0
    kmc.free()
