"""
The five benchmark configurations at their full sizes, on the device step path and
with the tile choices the benchmark uses, against the oracle:

cfg0  free space, N=1000, box 10, 8^3 leaves, L=12, 64 x 4 proposals (ball 1.0,
      clipped), Metropolis at kT=1, 100 steps: every step's dE against the oracle's
      extended reference (1e-10) and its reference-ordered FP64 restatement, the
      accepted sequence, positions and cell tables bit-exact, expansions within 1e-11.
cfg1  pbc, N=100000, box 100, 32^3 leaves, L=12, 1024 x 8 proposals (ball 1.5): 50
      steps (not 1000: the oracle's CPU accept sweep over 38^3 stored leaves takes
      ~0.3 s a step on the box's 16 host threads, and the whole GPU suite has to fit
      the 15-minute round-end run); dE of all 8192 proposals of every step against the
      extended reference, accept sequence, cell tables, positions bit-exact,
      expansions within 1e-11.
cfg2  the cfg1 system, mixed vs FP64 over 65536 proposals (8 steps of 8192): max /
      median absolute and floored relative error, mixed within 1e-8 and FP64 within
      1e-10 of the extended reference.
cfg3  256 systems (N log-uniform in [2000, 200000] at density 0.1, half free space,
      half pbc, L in {8, 12, 16}, 256 x 8 proposals + 1 accept each), batched kernels
      on 6 streams, mixed, 4 steps in captured graphs: a sample of systems re-run alone
      gives bit-identical selections, energies and expansions; no flags.
cfg4  pbc, N=2000000, box 271.4, 64^3 leaves, L=16: dE of 32768 proposals (4096 x 8,
      ball 1.5) of both paths against the extended reference; after one accept the
      expansions of sampled leaves and of the leaves around the move against the
      oracle's accept sweep (1e-11 of their scale), cell tables of the two touched
      leaves, and dE of the next proposals.
"""

import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import neutral_system, rel_err, metropolis_pick, move_stream, step_moves

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _print(*a):
    print(*a, flush=True)


def _dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


class DeviceStream:
    """The device-side move stream and step buffers of one system."""

    def __init__(self, kmc, pids, disp, us, P, k):
        import torch
        self.kmc, self.P, self.k, self.T = kmc, P, k, P * k
        self.t_pids = _dev(pids.reshape(-1), torch.int64)
        self.t_disp = _dev(disp.reshape(-1), torch.float64)
        self.t_u = _dev(us, torch.float64)
        self.counter = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.sel = torch.full((len(us),), -1, dtype=torch.int64, device="cuda")
        self.d_pid = torch.zeros(P, dtype=torch.int64, device="cuda")
        self.d_es = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
        self.d_np = torch.zeros(self.T * 3, dtype=torch.float64, device="cuda")
        self.d_dE = torch.zeros(self.T, dtype=torch.float64, device="cuda")
        # the device step allocates nothing: a first propose of this size sizes the buffers
        self.materialize()
        kmc.propose_device(self.d_pid, self.d_es, self.d_np, self.d_dE, precision="fp64")

    def materialize(self):
        self.kmc.materialize_moves(self.P, self.k, self.t_pids, self.t_disp, self.counter, self.d_pid, self.d_es,
                                   self.d_np)

    def step(self, precision="fp64"):
        """One device step (materialise, propose, select, accept); returns the step's
        new positions and dE."""
        import torch
        self.materialize()
        npos = self.d_np.cpu().numpy().reshape(-1, 3)
        self.kmc.step_device(self.d_pid, self.d_es, self.d_np, self.d_dE, self.t_u, 1.0, self.counter, self.sel,
                             precision=precision)
        torch.cuda.synchronize()
        return npos, self.d_dE.cpu().numpy().copy()


def _bench_tiles(kmc, L, T, ncps):
    """The tile choices bench.py uses for this class (coulomb_kmc_amd/tiles.json)."""
    from coulomb_kmc_amd import tune
    kmc.set_tiles(**tune.choose_tiles(L, T, ncps ** 3))


def _store_equal(kmc, orc, cells=None):
    occ_g, ids_g, p4_g = kmc.get_store()
    occ_c, ids_c, data_c = orc.get_store()
    assert np.array_equal(occ_g.astype(np.int64), occ_c)
    for c in (np.nonzero(occ_c)[0] if cells is None else cells):
        n = occ_c[c]
        assert np.array_equal(ids_g[c, :n].astype(np.int64), ids_c[c, :n])
        assert np.array_equal(p4_g[c, :n, :4], data_c[c, :n, :4])


def _run_steps(kmc, orc, E, bc, P, k, pids, disp, us, nsteps, tag, ext_every=1):
    """nsteps device steps against the oracle's propose + Metropolis pick + accept."""
    ds = DeviceStream(kmc, pids, disp, us, P, k)
    es = np.arange(P + 1, dtype=np.int64) * k
    worst_ext, worst_ref = 0.0, 0.0
    for s in range(nsteps):
        npos_g, dg = ds.step("fp64")
        npos_c = step_moves(orc.positions(), pids[s], disp[s], k, E, bc)
        assert np.array_equal(npos_g, npos_c), s
        dc = orc.propose_arrays(pids[s], es, npos_c)
        worst_ref = max(worst_ref, float(rel_err(dg, dc).max()))
        if s % ext_every == 0:
            ext = orc.propose_check(pids[s], es, npos_c, extended=True)
            e = float(rel_err(dg, ext).max())
            worst_ext = max(worst_ext, e)
            assert e <= O.DE_BAR_FP64, (tag, s, e)
        kc = metropolis_pick(dc, us[s])
        kg = int(ds.sel[s].item())
        assert kg == kc, (tag, s, kg, kc)
        orc.accept(int(pids[s][kc // k]), npos_c[kc])
    assert kmc.check_flags() == 0
    _print(tag, nsteps, "steps: dE vs extended reference max rel", worst_ext, "vs reference order", worst_ref)
    return ds


def test_cfg0_free_space_1k_100_steps():
    from coulomb_kmc_amd import KMCFMM
    N, E, r, L, P, k, R, nsteps = 1000, 10.0, 4, 12, 64, 4, 1.0, 100
    rng = np.random.default_rng(1000)
    pos, q = neutral_system(rng, N, E)
    pids, disp, us = move_stream(rng, N, P, k, R, nsteps)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition="free_space", max_move=R)
    _bench_tiles(kmc, L, P * k, 2 ** (r - 1))
    kmc.initialise()
    orc = O.Oracle(pos, q, E, r, L, bc="free_space", max_move=R)
    orc.initialise()
    assert abs(kmc.energy - orc.energy) <= 1e-11 * abs(orc.energy)
    _run_steps(kmc, orc, E, "free_space", P, k, pids, disp, us, nsteps, "cfg0")
    pos_g, cells_g = kmc.get_device_particles()
    assert np.array_equal(pos_g, orc.positions())
    assert np.array_equal(cells_g.astype(np.int64), orc.cells())
    _store_equal(kmc, orc)
    eg, ec = kmc.get_expansions(), orc.get_expansions()
    err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
    _print("cfg0 expansions after 100 accepts", err, "energy", kmc.device_energy(), orc.energy)
    assert err <= 1e-11
    assert abs(kmc.device_energy() - orc.energy) <= 1e-10 * abs(orc.energy)
    kmc.free()


@pytest.fixture(scope="module")
def pbc_100k():
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    N, E, r, L, R = 100000, 100.0, 6, 12, 1.5
    rng = np.random.default_rng(1001)
    pos, q = neutral_system(rng, N, E)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition="pbc", max_move=R)
    _bench_tiles(kmc, L, 1024 * 8, 2 ** (r - 1))
    kmc.initialise()
    orc = O.Oracle(pos, q, E, r, L, bc="pbc", max_move=R, lr=lr_operator_reference_layout(L, E))
    orc.init_from_expansions(kmc.get_expansions(), kmc.energy)
    yield dict(kmc=kmc, orc=orc, pos=pos, q=q, rng=rng, N=N, E=E, R=R)
    kmc.free()


def test_cfg2_mixed_vs_fp64_65536_proposals(pbc_100k):
    import torch
    s = pbc_100k
    kmc, orc, N, E = s["kmc"], s["orc"], s["N"], s["E"]
    P, k, nb = 1024, 8, 8
    pids, disp, _ = move_stream(np.random.default_rng(2002), N, P, k, s["R"], nb)
    es = np.arange(P + 1, dtype=np.int64) * k
    d_es = _dev(es, torch.int64)
    g64, gmx, ext = [], [], []
    for b in range(nb):
        npos = step_moves(s["pos"], pids[b], disp[b], k, E, "pbc")
        d_pid, d_np = _dev(pids[b], torch.int64), _dev(npos, torch.float64)
        out64 = torch.zeros(P * k, dtype=torch.float64, device="cuda")
        outmx = torch.zeros(P * k, dtype=torch.float64, device="cuda")
        kmc.propose_device(d_pid, d_es, d_np, out64, precision="fp64")
        kmc.propose_device(d_pid, d_es, d_np, outmx, precision="mixed")
        g64.append(out64.cpu().numpy())
        gmx.append(outmx.cpu().numpy())
        ext.append(orc.propose_check(pids[b], es, npos, extended=True))
    g64, gmx, ext = map(np.concatenate, (g64, gmx, ext))
    assert g64.size == 65536
    e_mix = O.de_error(gmx, g64)
    e_mix_ext = O.de_error(gmx, ext)
    e_64_ext = O.de_error(g64, ext)
    _print("cfg2 mixed vs fp64", e_mix)
    _print("cfg2 mixed vs extended", e_mix_ext)
    _print("cfg2 fp64 vs extended", e_64_ext)
    assert e_mix_ext["max_rel"] <= O.DE_BAR_MIXED
    assert e_64_ext["max_rel"] <= O.DE_BAR_FP64


def test_cfg1_pbc_100k_50_steps(pbc_100k):
    s = pbc_100k
    kmc, orc, N, E = s["kmc"], s["orc"], s["N"], s["E"]
    P, k, nsteps = 1024, 8, 50
    pids, disp, us = move_stream(np.random.default_rng(1011), N, P, k, s["R"], nsteps)
    _run_steps(kmc, orc, E, "pbc", P, k, pids, disp, us, nsteps, "cfg1")
    pos_g, cells_g = kmc.get_device_particles()
    assert np.array_equal(pos_g, orc.positions())
    assert np.array_equal(cells_g.astype(np.int64), orc.cells())
    _store_equal(kmc, orc)
    eg, ec = kmc.get_expansions(), orc.get_expansions()
    err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
    Mg = kmc.get_multipole()
    Mc, _ = orc.get_ME()
    _print("cfg1 expansions after 50 accepts", err, "multipole", np.max(np.abs(Mg - Mc)) / np.max(np.abs(Mc)))
    assert err <= 1e-11
    assert np.max(np.abs(Mg - Mc)) <= 1e-12 * np.max(np.abs(Mc))


def test_cfg3_batch_256_systems():
    import torch
    sys.path.insert(0, ROOT)
    import bench
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.batch import BatchSystem, KMCBatch
    spec = dict(bench.BATCH)
    specs, _ = bench.batch_systems(spec)
    assert len(specs) == 256
    nsteps, G = 4, 2
    streams_rng = np.random.default_rng(3003)
    moves = [move_stream(streams_rng, x["N"], x["P"], x["k"], x["radius"], nsteps) for x in specs]

    def make(i):
        x = specs[i]
        kmc = KMCFMM(x["pos"].copy(), x["q"].reshape(-1, 1), x["E"], r=x["r"], l=x["L"],
                     boundary_condition=x["bc"], max_move=x["radius"], precision="mixed", init_method="fmm")
        kmc.initialise()
        kmc.set_tiles(near_lanes=spec["near_lanes"], sweep_block=1, sweep_pairs=2, sweep_split=1)
        p, d, u = moves[i]
        return BatchSystem(kmc, x["P"], x["k"], _dev(p.reshape(-1), torch.int64), _dev(d.reshape(-1), torch.float64),
                           _dev(u, torch.float64))

    systems = [make(i) for i in range(len(specs))]
    batch = KMCBatch(systems, kT=1.0, precision="mixed", streams=6, mode="kernels")
    batch.capture(G)
    batch.replay(nsteps // G)
    torch.cuda.synchronize()
    flags = batch.check_flags()
    assert all(f == 0 for f in flags)
    # a sample over both boundary conditions, all three orders and the size range
    order = np.argsort([x["N"] for x in specs])
    sample = sorted(set([int(order[0]), int(order[-1]), int(order[len(order) // 2])] +
                        [next(i for i in range(256) if specs[i]["L"] == L and specs[i]["bc"] == bc)
                         for L in (8, 12, 16) for bc in ("free_space", "pbc")]))
    for i in sample:
        one = make(i)
        one.materialize()
        one.kmc.propose_device(one.d_pid, one.d_es, one.d_new, one.d_dE, precision="mixed")
        for _ in range(nsteps):
            one.step(1.0, "mixed")
        torch.cuda.synchronize()
        sb = systems[i]
        _print("cfg3 system", i, specs[i]["N"], specs[i]["bc"], specs[i]["L"], sb.selected.cpu().numpy()[:nsteps])
        assert np.array_equal(sb.selected.cpu().numpy(), one.selected.cpu().numpy())
        assert sb.kmc.device_energy() == one.kmc.device_energy()
        assert np.array_equal(sb.kmc.get_expansions(), one.kmc.get_expansions())
        assert one.kmc.check_flags() == 0
        one.kmc.free()
    batch.free()
    for s in systems:
        s.kmc.free()


def test_cfg4_pbc_2m():
    import torch
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    N, E, r, L, R = 2000000, 271.4, 7, 16, 1.5
    rng = np.random.default_rng(1004)
    pos, q = neutral_system(rng, N, E)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition="pbc", max_move=R)
    _bench_tiles(kmc, L, 32768 * 8, 2 ** (r - 1))
    kmc.initialise()
    orc = O.Oracle(pos, q, E, r, L, bc="pbc", max_move=R, lr=lr_operator_reference_layout(L, E))
    orc.init_from_expansions(kmc.get_expansions(), kmc.energy)
    P, k = 4096, 8
    pids, disp, _ = move_stream(np.random.default_rng(4004), N, P, k, R, 2)
    es = np.arange(P + 1, dtype=np.int64) * k
    d_es = _dev(es, torch.int64)

    def both_paths(s, cur):
        npos = step_moves(cur, pids[s], disp[s], k, E, "pbc")
        d_pid, d_np = _dev(pids[s], torch.int64), _dev(npos, torch.float64)
        out = {}
        for prec in ("fp64", "mixed"):
            t = torch.zeros(P * k, dtype=torch.float64, device="cuda")
            kmc.propose_device(d_pid, d_es, d_np, t, precision=prec)
            out[prec] = t.cpu().numpy()
        ext = orc.propose_check(pids[s], es, npos, extended=True)
        return npos, out, ext

    npos, out, ext = both_paths(0, pos)
    e64, emx = O.de_error(out["fp64"], ext), O.de_error(out["mixed"], ext)
    _print("cfg4 fp64 vs extended", e64)
    _print("cfg4 mixed vs extended", emx)
    assert e64["max_rel"] <= O.DE_BAR_FP64 and emx["max_rel"] <= O.DE_BAR_MIXED
    # one accept: proposal 0's move, on both sides
    pid, newp = int(pids[0][0]), npos[0]
    kmc.accept((pid, newp))
    orc.accept(pid, newp)
    ncps = 2 ** (r - 1)
    c3o = np.minimum(((pos[pid] + 0.5 * E) / (E / ncps)).astype(int), ncps - 1)
    c3n = np.minimum(((newp + 0.5 * E) / (E / ncps)).astype(int), ncps - 1)
    lin = lambda c: int(c[0] + ncps * (c[1] + ncps * c[2]))
    eg, ec = kmc.get_expansions(), orc.get_expansions()
    leaves = np.unique(np.concatenate([np.random.default_rng(5).choice(ncps ** 3, 4096, replace=False),
                                       [lin((c3o + d) % ncps) for d in np.ndindex(5, 5, 5) for d in [np.array(d) - 2]],
                                       [lin((c3n + d) % ncps) for d in np.ndindex(5, 5, 5) for d in [np.array(d) - 2]]]))
    err = np.max(np.abs(eg[leaves] - ec[leaves])) / np.max(np.abs(ec[leaves]))
    _print("cfg4 expansions after one accept,", len(leaves), "leaves:", err)
    assert err <= 1e-11
    del eg, ec
    occ_g, ids_g, p4_g = kmc.get_store()
    occ_c, ids_c, data_c = orc.get_store()
    for c in (lin(c3o), lin(c3n)):
        n = occ_c[c]
        assert occ_g[c] == n
        assert np.array_equal(ids_g[c, :n].astype(np.int64), ids_c[c, :n])
        assert np.array_equal(p4_g[c, :n, :4], data_c[c, :n, :4])
    cur = pos.copy()
    cur[pid] = newp
    npos, out, ext = both_paths(1, cur)
    e64 = O.de_error(out["fp64"], ext)
    _print("cfg4 after accept fp64 vs extended", e64)
    assert e64["max_rel"] <= O.DE_BAR_FP64
    kmc.free()
