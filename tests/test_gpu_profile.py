"""The per-kernel device profile (kmc_profile_enable / kmc_profile_read; the reference's
c_direct_* / c_indirect_* / c_accept_* counters, coulomb_kmc/kmc_local.py:466-492,
kmc_octal.py:176-183, 352-358): evaluation-point, sweep and near-field pair counts
exact against a host count over the cell tables; times positive; nothing recorded
while off or under graph capture."""

import numpy as np
import pytest

from helpers import neutral_system, ball_moves

pytestmark = pytest.mark.gpu


def _pairs(kmc, pts):
    """Sum over the points of the occupancies of the 27 leaves around each point's leaf."""
    occ = kmc.get_store()[0]
    n = round(kmc.ncells ** (1.0 / 3.0))
    tot = 0
    for tx, ty, tz in kmc.leaf_of(pts):   # leaf index per axis
        for ox in (-1, 0, 1):
            for oy in (-1, 0, 1):
                for oz in (-1, 0, 1):
                    x, y, z = tx + ox, ty + oy, tz + oz
                    if kmc.boundary_condition.value == "free_space":
                        if min(x, y, z) < 0 or max(x, y, z) >= n:
                            continue
                    tot += int(occ[(x % n) + n * ((y % n) + n * (z % n))])
    return tot


@pytest.mark.parametrize("bc", ["pbc", "free_space"])
def test_kernel_profile_counts(bc):
    import torch
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(5)
    N, E = 3000, 16.0
    pos, q = neutral_system(rng, N, E)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=4, l=12, boundary_condition=bc, max_move=1.0)
    kmc.initialise()
    kmc.profile_kernels(True)
    want_pts, want_pairs, want_lr = 0, 0, 0
    for _ in range(3):
        pids = rng.choice(N, 40, replace=False)
        moves = ball_moves(rng, kmc._pos(), pids, 5, 1.0, E, bc)
        kmc.propose(moves)
        newp = np.concatenate([m[1] for m in moves])
        oldp = kmc._pos()[pids]
        want_pts += len(newp) + len(pids)
        want_pairs += _pairs(kmc, newp) + _pairs(kmc, oldp)
        want_lr += len(newp)
    prof = kmc.kernel_profile()
    assert prof["c_direct_exec_count"] == want_pts == prof["c_indirect_exec_count"]
    assert prof["c_direct_div_count"] == want_pairs
    assert prof["c_long_range_exec_count"] == (want_lr if bc == "pbc" else 0)
    assert prof["c_direct_time"] > 0 and prof["c_indirect_time"] > 0
    assert prof["c_accept_exec_count"] == 0
    for _ in range(2):
        pid = int(rng.integers(N))
        kmc.accept((pid, ball_moves(rng, kmc._pos(), [pid], 1, 1.0, E, bc)[0][1][0]))
    prof = kmc.kernel_profile()
    assert prof["c_accept_exec_count"] == 2 and prof["c_accept_time"] > 0
    assert prof["c_accept_gflop_rate"] > 0 and prof["c_direct_div_sqrt_grate"] > 0
    assert kmc.profile["c_accept_exec_count"] == 2
    # off: nothing more is recorded (and the sums are cleared)
    kmc.profile_kernels(False)
    pids = rng.choice(N, 10, replace=False)
    kmc.propose(ball_moves(rng, kmc._pos(), pids, 3, 1.0, E, bc))
    prof = kmc.kernel_profile()
    assert prof["c_direct_exec_count"] == 0 and prof["c_accept_exec_count"] == 0
    # on, but under graph capture: no event is recorded
    kmc.profile_kernels(True)
    P, k = 8, 3
    d = rng.normal(size=(P * k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
    t_pids = dev(rng.choice(N, P, replace=False), torch.int64)
    t_disp = dev(0.5 * d.reshape(-1), torch.float64)
    us = dev(rng.uniform(size=4), torch.float64)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    sel = torch.full((4,), -1, dtype=torch.int64, device="cuda")
    d_pid = torch.zeros(P, dtype=torch.int64, device="cuda")
    d_es = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
    d_new = torch.zeros(P * k * 3, dtype=torch.float64, device="cuda")
    d_dE = torch.zeros(P * k, dtype=torch.float64, device="cuda")
    kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_new)
    kmc.propose_device(d_pid, d_es, d_new, d_dE)     # eager: recorded
    torch.cuda.synchronize()
    before = kmc.kernel_profile()
    assert before["c_direct_exec_count"] == P * k + P
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_new)
        kmc.step_device(d_pid, d_es, d_new, d_dE, us, 1.0, counter, sel)
    g.replay()
    torch.cuda.synchronize()
    after = kmc.kernel_profile()
    assert after["c_direct_exec_count"] == before["c_direct_exec_count"]
    assert after["c_accept_exec_count"] == before["c_accept_exec_count"]
    kmc.free()
