"""
Pins the order truncation of the blocked accept sweep on the CPU (no GPU):

The blocked sweep skips orders l >= lc of a source image whose nearest leaf centre in
the 4x4x4 block is at distance D, with lc the smallest order such that
|q| (rho/D)^lc / (D - rho) <= 2^-60 |q| (rho: a leaf's half diagonal).  By the addition
theorem the skipped orders change the potential the leaf's local expansion gives at
any point of the leaf by at most that bound.  Checked here with the oracle's own P2L
and local evaluation (the reference's basis, kmc_expansion_tools.py:154-226): the
potential of the truncated expansion differs from the full one by no more than the
bound, and the rule is the one the kernel applies (same lc for a range of distances).
"""

import numpy as np

import oracle as O

TAU = 2.0 ** -60


def trunc_order(D, rho, L):
    """The kernel's rule (sweep_trunc_order in csrc/kmc_device.h)."""
    if not D > 3.0 * rho:
        return L
    t = rho / D
    bound = 1.0 / (D - rho)
    lc = 0
    while lc < L and bound > TAU:
        bound *= t
        lc += 1
    return lc


def truncate(lexp, L, lc):
    """Zero the orders l >= lc of a reference-layout expansion [2 L^2]."""
    out = np.array(lexp, copy=True)
    for l in range(lc, L):
        for m in range(-l, l + 1):
            out[l * l + l + m] = 0.0
            out[L * L + l * l + l + m] = 0.0
    return out


def test_truncation_bound_holds():
    rng = np.random.default_rng(17)
    L, w = 16, 1.0
    rho = 0.5 * np.sqrt(3.0) * w
    seen = set()
    for D in (2.5, 4.0, 8.0, 20.0, 60.0, 150.0, 400.0):
        lc = trunc_order(D, rho, L)
        seen.add(lc)
        for _ in range(6):
            d = rng.normal(size=3)
            src = D * d / np.linalg.norm(d)               # source at distance D from the centre
            q = rng.choice([-1.0, 1.0])
            full = O.p2l(L, q, src)
            trunc = truncate(full, L, lc)
            bound = abs(q) * (rho / D) ** lc / (D - rho)
            for _ in range(8):
                x = rng.uniform(-0.5 * w, 0.5 * w, 3)    # a point in the leaf
                dphi = O.local_eval(L, full, x) - O.local_eval(L, trunc, x)
                # FP64 evaluation of O(1/D) terms adds ~1e-16/D of rounding on top
                assert abs(dphi) <= bound + 4e-16 / D, (D, lc, dphi, bound)
    # the rule truncates far images and keeps every order of near ones
    assert trunc_order(2.5, rho, L) == L
    assert min(seen) < 10


def test_truncation_rule_values():
    rho = 0.5 * np.sqrt(3.0) * 271.4 / 64
    # bound at lc is below 2^-60, at lc - 1 above (minimality), for the 2M-charge leaf size
    for D in (15.0, 40.0, 100.0, 200.0, 400.0):
        lc = trunc_order(D, rho, 16)
        if lc < 16:
            assert (rho / D) ** lc / (D - rho) <= TAU
        if lc > 0:
            assert (rho / D) ** (lc - 1) / (D - rho) > TAU


def test_near_leaves_keep_every_order():
    """Truncated images skip the near-leaf test: a leaf of the block whose 27-leaf
    stencil holds the image's leaf lies within 1.5 leaf widths of the image on each
    axis, so the image's distance D to the box of the block's leaf centres is below
    3 rho, where the rule keeps every order.  Brute force over random images and
    blocks, with the kernel's leaf binning (floor) and its block box."""
    rng = np.random.default_rng(23)
    w, ncps = 1.7, 16
    rho = 0.5 * np.sqrt(3.0) * w
    lo = -0.5 * ncps * w
    worst = 0.0
    for _ in range(4000):
        b = rng.integers(0, ncps // 4, 3)
        mid = lo + (4 * b + 2.0) * w                                   # centre of the block's leaf-centre box
        p = mid + rng.uniform(-4.0 * w, 4.0 * w, 3)                    # images around the block
        cell = np.floor((p - lo) / w).astype(int)
        leaves = 4 * b + np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3)
        near = np.all(np.abs(cell - leaves) <= 1, axis=1)
        e = np.maximum(np.abs(p - mid) - 1.5 * w, 0.0)
        D = float(np.sqrt((e ** 2).sum()))
        if near.any():
            worst = max(worst, D / rho)
            assert D < 3.0 * rho, (p, b, D / rho)
            assert trunc_order(D, rho, 16) == 16
    assert worst > 2.0    # the cases seen come close to the 3 rho limit

