"""kmc_propose_noalloc (the propose entry point that allocates nothing: its evaluation
arrays live in the library's fixed device scratch array) against kmc_propose, called
through the C ABI: two calls in a row, the second smaller, results within the parity
bar (bit-identical in practice: same kernels, same order of operations); oversized
batches and bad moves are refused as documented.  Named to run last in the suite."""

import ctypes

import numpy as np
import pytest

from helpers import neutral_system, ball_moves, rel_err

pytestmark = pytest.mark.gpu


def _system(bc, N, E, r, L, seed):
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(seed)
    pos, q = neutral_system(rng, N, E)
    kmc = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=1.0)
    kmc.initialise()
    return kmc, pos, rng


def _moves(rng, pos, N, P, k, E, bc):
    import torch
    pids = rng.choice(N, P, replace=False)
    mv = ball_moves(rng, pos, pids, k, 1.0, E, bc)
    npos = np.concatenate([m[1] for m in mv])
    es = np.arange(P + 1, dtype=np.int64) * k
    dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
    return dev(pids, torch.int64), dev(es, torch.int64), dev(npos, torch.float64), P * k


def _call(fn, kmc, d_pid, d_es, d_np, T, prec):
    import torch
    out = torch.full((T,), float("nan"), dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(kmc.handle, st, d_pid.numel(), T, d_pid.data_ptr(), d_es.data_ptr(), d_np.data_ptr(), None,
            out.data_ptr(), prec)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy()


@pytest.mark.parametrize("bc,prec", [("pbc", 0), ("pbc", 1), ("free_space", 0)])
def test_noalloc_equals_propose_twice(bc, prec):
    from coulomb_kmc_amd import _native
    lib = _native.lib()
    N, E = 4000, 20.0
    kmc, pos, rng = _system(bc, N, E, 4, 12, 31)
    for P, k in ((512, 8), (96, 4)):     # second call smaller
        d_pid, d_es, d_np, T = _moves(rng, pos, N, P, k, E, bc)
        print("noalloc", bc, prec, "P", P, "T", T, flush=True)
        rc0, ref = _call(lib.kmc_propose, kmc, d_pid, d_es, d_np, T, prec)
        rc1, got = _call(lib.kmc_propose_noalloc, kmc, d_pid, d_es, d_np, T, prec)
        assert rc0 == 0 and rc1 == 0, _native.lib().kmc_last_error()
        assert np.all(np.isfinite(got))
        assert rel_err(got, ref).max() <= 1e-10
        assert np.array_equal(got, ref)
    kmc.free()


def test_noalloc_rejects_oversize_and_bad_moves():
    import torch
    from coulomb_kmc_amd import _native
    lib = _native.lib()
    N, E = 4000, 20.0
    kmc, pos, rng = _system("pbc", N, E, 4, 12, 32)
    # more proposals than the fixed scratch holds: refused before any launch
    big_T = 1 << 21
    d_pid = torch.zeros(1, dtype=torch.int64, device="cuda")
    d_es = torch.tensor([0, big_T], dtype=torch.int64, device="cuda")
    d_np = torch.zeros(3 * big_T, dtype=torch.float64, device="cuda")
    rc, _ = _call(lib.kmc_propose_noalloc, kmc, d_pid, d_es, d_np, big_T, 0)
    assert rc == _native.KMC_E_ARG
    assert b"scratch" in lib.kmc_last_error()
    # a move beyond max_move: the same error as kmc_propose
    d_pid = torch.tensor([5], dtype=torch.int64, device="cuda")
    d_es = torch.tensor([0, 1], dtype=torch.int64, device="cuda")
    far = pos[5] + np.array([3.0, 0.0, 0.0])
    far = np.where(far >= 0.5 * E, far - E, far)
    d_np = torch.as_tensor(far, dtype=torch.float64).cuda()
    rc0, _ = _call(lib.kmc_propose, kmc, d_pid, d_es, d_np, 1, 0)
    rc1, _ = _call(lib.kmc_propose_noalloc, kmc, d_pid, d_es, d_np, 1, 0)
    assert rc0 == rc1 == _native.KMC_E_BAD_MOVE
    kmc.free()
