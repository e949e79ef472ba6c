"""kmc_run_steps: runs of materialise + step in one call, with the accept sweep of each
step overlapping the next step's set-up, near field and long range (pipelined), give
exactly the results of calling kmc_materialize_moves and kmc_step one step at a time:
same selected proposals, energies, positions and leaf expansions, eager and replayed
from a captured graph."""

import numpy as np
import pytest

from helpers import neutral_system

pytestmark = pytest.mark.gpu


def _setup(bc, seed=11, N=1500, E=10.0, r=4, L=12, P=24, k=4, nst=7):
    import torch
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(seed)
    pos, q = neutral_system(rng, N, E)
    pids = np.stack([rng.choice(N, P, replace=False) for _ in range(nst)])
    d = rng.normal(size=(nst, P * k, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    d *= rng.uniform(size=(nst, P * k, 1)) ** (1 / 3)
    us = rng.uniform(size=nst)
    dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
    tabs = dict(pids=dev(pids.reshape(-1), torch.int64), disp=dev(d.reshape(-1), torch.float64),
                us=dev(us, torch.float64))

    def make():
        kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=1.0,
                     precision="mixed")
        kmc.initialise()
        w = dict(counter=torch.zeros(1, dtype=torch.int64, device="cuda"),
                 sel=torch.full((nst,), -1, dtype=torch.int64, device="cuda"),
                 pid=torch.zeros(P, dtype=torch.int64, device="cuda"),
                 es=torch.zeros(P + 1, dtype=torch.int64, device="cuda"),
                 new=torch.zeros(P * k * 3, dtype=torch.float64, device="cuda"),
                 dE=torch.zeros(P * k, dtype=torch.float64, device="cuda"))
        kmc.materialize_moves(P, k, tabs["pids"], tabs["disp"], w["counter"], w["pid"], w["es"], w["new"])
        kmc.propose_device(w["pid"], w["es"], w["new"], w["dE"])   # sizes the work buffers
        torch.cuda.synchronize()
        return kmc, w
    return make, tabs, P, k


def _state(kmc, w):
    import torch
    torch.cuda.synchronize()
    pos, _ = kmc.get_device_particles()
    lexp = kmc.get_expansions() if kmc.boundary_condition.value != "ff-only" else None
    return (int(w["counter"].item()), w["sel"].cpu().numpy(), kmc.device_energy(), pos, lexp)


def _same(a, b):
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]
    assert np.array_equal(a[3], b[3])
    assert (a[4] is None and b[4] is None) or np.array_equal(a[4], b[4])


@pytest.mark.parametrize("bc", ["pbc", "27", "free_space", "ff-only"])
def test_run_steps_equal_single_steps(bc):
    import torch
    make, tabs, P, k = _setup(bc)
    nrun = 6
    ref_k, ref_w = make()
    for _ in range(nrun):
        ref_k.materialize_moves(P, k, tabs["pids"], tabs["disp"], ref_w["counter"], ref_w["pid"], ref_w["es"],
                                ref_w["new"])
        ref_k.step_device(ref_w["pid"], ref_w["es"], ref_w["new"], ref_w["dE"], tabs["us"], 1.0, ref_w["counter"],
                          ref_w["sel"])
    ref = _state(ref_k, ref_w)
    assert ref[0] == nrun
    for pipelined in (0, 1, 2, 3, 4):
        kk, w = make()
        kk.run_steps(nrun, P, k, tabs["pids"], tabs["disp"], tabs["us"], 1.0, w["counter"], w["sel"], w["pid"],
                     w["es"], w["new"], w["dE"], pipelined=pipelined)
        _same(_state(kk, w), ref)
        kk.free()
    # captured (every pipelined mode): 3 steps per graph, replayed twice with eager work
    # between the replays
    for pipelined in (1, 2, 3, 4):
        kk, w = make()
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            kk.run_steps(3, P, k, tabs["pids"], tabs["disp"], tabs["us"], 1.0, w["counter"], w["sel"], w["pid"],
                         w["es"], w["new"], w["dE"], pipelined=pipelined, stream=s.cuda_stream)
        for _ in range(2):
            g.replay()
            torch.cuda.synchronize()
            _ = w["counter"].cpu()
        _same(_state(kk, w), ref)
        del g
        kk.free()
    ref_k.free()


def test_run_steps_rejects_unsized_buffers():
    import torch
    make, tabs, P, k = _setup("pbc", N=600, nst=3)
    kk, w = make()
    with pytest.raises(RuntimeError):
        # twice the proposals of the sized buffers
        kk.run_steps(1, 2 * P, k, tabs["pids"][:2 * P], torch.zeros(2 * P * k * 3, dtype=torch.float64,
                                                                         device="cuda"),
                     tabs["us"], 1.0, w["counter"], w["sel"], torch.zeros(2 * P, dtype=torch.int64, device="cuda"),
                     torch.zeros(2 * P + 1, dtype=torch.int64, device="cuda"),
                     torch.zeros(2 * P * k * 3, dtype=torch.float64, device="cuda"),
                     torch.zeros(2 * P * k, dtype=torch.float64, device="cuda"))
    kk.free()
