"""Batched stepping (coulomb_kmc_amd/batch.py): several independent systems of different
orders and boundary conditions, stepped by the batched kernels (one launch per kernel
kind, system = blockIdx.z) or by their own steps in one graph, give exactly the results
of stepping each system alone (same move streams, same tiles)."""

import numpy as np
import pytest

from helpers import neutral_system

pytestmark = pytest.mark.gpu


def _systems(seed):
    import torch
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.batch import BatchSystem
    rng = np.random.default_rng(seed)
    out = []
    for i, (N, E, r, L, bc) in enumerate(((600, 8.0, 4, 8, "pbc"), (900, 10.0, 4, 12, "free_space"),
                                           (1500, 12.0, 5, 8, "27"), (700, 8.0, 4, 12, "pbc"),
                                           (800, 9.0, 4, 12, "ff-only"))):
        pos, q = neutral_system(rng, N, E)
        k = KMCFMM(pos, q, E, r=r, l=L, boundary_condition=bc, max_move=1.0, precision="mixed")
        k.initialise()
        P, kk, nst = 16, 4, 6
        pids = np.stack([rng.choice(N, P, replace=False) for _ in range(nst)])
        d = rng.normal(size=(nst, P * kk, 3))
        d /= np.linalg.norm(d, axis=2, keepdims=True)
        d *= rng.uniform(size=(nst, P * kk, 1)) ** (1 / 3)
        us = rng.uniform(size=nst)
        dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
        out.append(BatchSystem(k, P, kk, dev(pids.reshape(-1), torch.int64), dev(d.reshape(-1), torch.float64),
                               dev(us, torch.float64)))
    return out


def _check_batch(streams, mode="graph", near_lanes=27, sweep_block=0):
    import torch
    from coulomb_kmc_amd.batch import KMCBatch
    steps = 4
    batched = _systems(5)
    single = _systems(5)
    for s in batched + single:   # the batched kernels' shapes
        s.kmc.set_tiles(near_lanes=near_lanes, sweep_block=sweep_block, sweep_pairs=2, sweep_split=1)
    b = KMCBatch(batched, kT=1.0, precision="mixed", streams=streams, mode=mode)
    b.capture(steps_per_graph=2)
    b.replay(steps // 2)
    torch.cuda.synchronize()
    for s in single:
        s.materialize()
        s.kmc.propose_device(s.d_pid, s.d_es, s.d_new, s.d_dE, precision="mixed")
        for _ in range(steps):
            s.step(1.0, "mixed")
    torch.cuda.synchronize()
    for sb, ss in zip(batched, single):
        assert int(sb.counter.item()) == steps == int(ss.counter.item())
        assert np.array_equal(sb.selected.cpu().numpy(), ss.selected.cpu().numpy())
        assert sb.kmc.device_energy() == ss.kmc.device_energy()
        if sb.kmc.boundary_condition.value != "ff-only":
            assert np.array_equal(sb.kmc.get_expansions(), ss.kmc.get_expansions())
    assert all(f == 0 for f in b.check_flags())
    b.free()
    for s in batched + single:
        s.kmc.free()


def test_batch_graph_equals_single_system_steps():
    _check_batch(1, "graph")


def test_batched_kernels_equal_single_system_steps():
    _check_batch(1, "kernels")


@pytest.mark.parametrize("near_lanes", [4, 8, 16])
def test_batched_per_point_near_equals_single_system_steps(near_lanes):
    # every system selects the same per-point lane count: the batch uses kb_near<G>
    _check_batch(1, "kernels", near_lanes)


@pytest.mark.parametrize("streams", [1, 3])
def test_batched_blocked_sweep_equals_single_system_steps(streams):
    # sweep_block 1 everywhere: one blocked-sweep launch per group over its flat
    # (system, 4x4x4 block) table
    _check_batch(streams, "kernels", 8, sweep_block=1)


@pytest.mark.skip(reason="recording the systems on extra torch side streams crashed the host process inside "
                         "graph instantiation (torch.cuda.graph capture_end) on this ROCm/PyTorch stack; "
                         "KMCBatch raises for streams != 1")
def test_batch_graph_on_side_streams():
    _check_batch(3, "graph")


def test_batch_rejects_side_streams():
    from coulomb_kmc_amd.batch import KMCBatch
    with pytest.raises(NotImplementedError):
        KMCBatch([], streams=2, mode="graph")
    with pytest.raises(ValueError):
        KMCBatch([], streams=17)


@pytest.mark.parametrize("streams", [2, 4])
def test_batched_kernels_on_side_streams_equal_single_system_steps(streams):
    # the per-group launches on library side streams (event fork / join inside the step)
    _check_batch(streams, "kernels", 8)
