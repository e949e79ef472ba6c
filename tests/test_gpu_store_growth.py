"""Cell-table growth for device-side steps (kmc_store_reserve; the reference grows its
particle store inside accept when the target cell is full, coulomb_kmc/kmc_local.py:
321-330, 558-578).  Every leaf starts full (the same number of charges in each leaf,
occupancy headroom 0): a run without reserving room drops selected moves (flag 2), and
runs that reserve room before each chunk of steps grow the tables and give exactly the
steps of a context created with ample headroom (same selections, energies, positions,
cell tables and expansions), one context at a time and in a batch replayed from a graph."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NPL = 4   # charges per leaf


def _full_leaves(rng, E, ncps):
    w = E / ncps
    pos = []
    for cz in range(ncps):
        for cy in range(ncps):
            for cx in range(ncps):
                lo = -0.5 * E + w * np.array([cx, cy, cz])
                pos.append(lo + w * rng.uniform(0.1, 0.9, (NPL, 3)))
    pos = np.concatenate(pos)
    q = np.ones(len(pos))
    q[: len(pos) // 2] = -1.0
    rng.shuffle(q)
    return pos, q


def _setup(bc, seed, P=16, k=4, nst=12, E=8.0, r=3, L=8):
    import torch
    rng = np.random.default_rng(seed)
    ncps = 2 ** (r - 1)
    pos, q = _full_leaves(rng, E, ncps)
    N = len(pos)
    pids = np.stack([rng.choice(N, P, replace=False) for _ in range(nst)])
    d = rng.normal(size=(nst, P * k, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    d *= rng.uniform(size=(nst, P * k, 1)) ** (1 / 3)
    us = rng.uniform(size=nst)
    dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
    tabs = dict(pids=dev(pids.reshape(-1), torch.int64), disp=dev(d.reshape(-1), torch.float64),
                us=dev(us, torch.float64))

    def make(headroom):
        from coulomb_kmc_amd import KMCFMM
        kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=1.0,
                     precision="mixed", occ_headroom=headroom)
        kmc.initialise()
        return kmc
    return make, tabs, P, k, nst


def _work(kmc, tabs, P, k, nst):
    import torch
    w = dict(counter=torch.zeros(1, dtype=torch.int64, device="cuda"),
             sel=torch.full((nst,), -1, dtype=torch.int64, device="cuda"),
             pid=torch.zeros(P, dtype=torch.int64, device="cuda"),
             es=torch.zeros(P + 1, dtype=torch.int64, device="cuda"),
             new=torch.zeros(P * k * 3, dtype=torch.float64, device="cuda"),
             dE=torch.zeros(P * k, dtype=torch.float64, device="cuda"))
    kmc.materialize_moves(P, k, tabs["pids"], tabs["disp"], w["counter"], w["pid"], w["es"], w["new"])
    kmc.propose_device(w["pid"], w["es"], w["new"], w["dE"])   # sizes the work buffers
    torch.cuda.synchronize()
    return w


def _run(kmc, w, tabs, P, k, n):
    kmc.run_steps(n, P, k, tabs["pids"], tabs["disp"], tabs["us"], 1.0, w["counter"], w["sel"], w["pid"], w["es"],
                  w["new"], w["dE"], pipelined=0)


def _tables(kmc):
    occ, ids, p4 = kmc.get_store()
    return occ, [ids[c, :o].copy() for c, o in enumerate(occ)], [p4[c, :o].copy() for c, o in enumerate(occ)]


def _same_state(a, b, wa, wb):
    import torch
    torch.cuda.synchronize()
    assert int(wa["counter"].item()) == int(wb["counter"].item())
    assert np.array_equal(wa["sel"].cpu().numpy(), wb["sel"].cpu().numpy())
    assert a.device_energy() == b.device_energy()
    assert np.array_equal(a.get_device_particles()[0], b.get_device_particles()[0])
    ta, tb = _tables(a), _tables(b)
    assert np.array_equal(ta[0], tb[0])
    assert all(np.array_equal(x, y) for x, y in zip(ta[1], tb[1]))
    assert all(np.array_equal(x, y) for x, y in zip(ta[2], tb[2]))
    assert np.array_equal(a.get_expansions(), b.get_expansions())


@pytest.mark.parametrize("bc", ["pbc", "free_space"])
def test_reserve_grows_store_and_matches_ample_headroom(bc):
    make, tabs, P, k, nst = _setup(bc, 31)
    # a context whose leaves are all full drops the selected moves into other leaves
    full = make(0)
    assert full.get_store()[1].shape[1] == NPL
    wf = _work(full, tabs, P, k, nst)
    _run(full, wf, tabs, P, k, nst)
    assert full.check_flags() & 2
    full.free()

    ref = make(64)
    wr = _work(ref, tabs, P, k, nst)
    _run(ref, wr, tabs, P, k, nst)
    assert ref.check_flags() == 0

    grown = make(0)
    wg = _work(grown, tabs, P, k, nst)
    grew = []
    for _ in range(nst // 3):
        grew.append(grown.reserve_store(3))
        _run(grown, wg, tabs, P, k, 3)
    assert grew[0]                       # every leaf was full
    assert grown.check_flags() == 0
    assert grown.get_store()[1].shape[1] > NPL
    _same_state(grown, ref, wg, wr)
    # nothing left to reserve for: no growth
    assert not grown.reserve_store(0)
    ref.free()
    grown.free()


def test_batch_replay_with_reserve_matches_ample_headroom():
    import torch
    from coulomb_kmc_amd.batch import BatchSystem, KMCBatch
    runs = {}
    for headroom in (0, 64):
        systems = []
        for i, bc in enumerate(("pbc", "free_space", "pbc")):
            make, tabs, P, k, nst = _setup(bc, 40 + i)
            kmc = make(headroom)
            kmc.set_tiles(near_lanes=8, sweep_block=1, sweep_pairs=2, sweep_split=1)
            systems.append(BatchSystem(kmc, P, k, tabs["pids"], tabs["disp"], tabs["us"]))
        b = KMCBatch(systems, kT=1.0, precision="mixed", streams=1, mode="kernels")
        b.capture(2)
        b.replay(4, reserve=(headroom == 0))
        torch.cuda.synchronize()
        assert all(f == 0 for f in b.check_flags())
        runs[headroom] = (b, systems)
    for sa, sb in zip(runs[0][1], runs[64][1]):
        assert int(sa.counter.item()) == 8 == int(sb.counter.item())
        assert np.array_equal(sa.selected.cpu().numpy(), sb.selected.cpu().numpy())
        assert sa.kmc.device_energy() == sb.kmc.device_energy()
        assert np.array_equal(sa.kmc.get_device_particles()[0], sb.kmc.get_device_particles()[0])
        assert np.array_equal(sa.kmc.get_expansions(), sb.kmc.get_expansions())
        assert sa.kmc.get_store()[1].shape[1] > NPL
    for b, systems in runs.values():
        b.free()
        for s in systems:
            s.kmc.free()
