"""The blocked accept sweep reads and writes only the orders below the bucket of each
4x4x4 block's largest truncation order (the higher orders receive no term from any
image).  After one hop every entry of those higher orders is bit-identical to its
value before the hop (a cut placed too low, or a wrong index in the update with its
run-time entry count, would lose or corrupt terms), and every expansion agrees with
the fused layout's (every order of every image, no truncation) to rounding.  That the
skipped entries are not read or written is what the HBM counters show (FETCH_SIZE +
WRITE_SIZE per launch against the algorithmic bytes).  The boxes are large (free space: 10^4 per
side on 16^3 leaves; periodic, where an image of each source is within half a box of
every block: 10^5 per side on 32^3 leaves; L = 16, and 15 for the wave pairs with a
one-column middle pair) so that the bound
(rho/D)^lc / (D - rho) <= 2^-60, which is absolute, keeps 12 or fewer orders for many
blocks; the test checks that such blocks exist."""

import numpy as np
import pytest

from test_sweep_truncation import trunc_order

pytestmark = pytest.mark.gpu

BUCKETS = (8, 9, 10, 12)


def bucket_of(lc, L):
    """Orders the blocked sweep evaluates for an image of truncation order lc (0: none)."""
    if lc <= 0:
        return 0
    for b in BUCKETS:
        if b < L and lc <= b:
            return b
    return L


def block_orders(E, ncps, L, periodic, old, new):
    """Per leaf (index x + n (y + n z)): the bucket of the largest truncation order over
    the images of both sources against the leaf's 4x4x4 block."""
    w = E / ncps
    rho = 0.5 * np.sqrt(3.0) * w
    c0 = -0.5 * E + 0.5 * w
    nb = ncps // 4
    R = 1 if periodic else 0
    out = np.zeros(ncps ** 3, dtype=int)
    for bz in range(nb):
        for by in range(nb):
            for bx in range(nb):
                mid = c0 + (4 * np.array([bx, by, bz]) + 1.5) * w
                lc = 0
                for src in (old, new):
                    for oz in range(-R, R + 1):
                        for oy in range(-R, R + 1):
                            for ox in range(-R, R + 1):
                                s = np.asarray(src) + E * np.array([ox, oy, oz])
                                e = np.maximum(np.abs(s - mid) - 1.5 * w, 0.0)
                                lc = max(lc, trunc_order(float(np.sqrt((e ** 2).sum())), rho, L))
                b = bucket_of(lc, L)
                for z in range(4):
                    for y in range(4):
                        for x in range(4):
                            out[(4 * bx + x) + ncps * ((4 * by + y) + ncps * (4 * bz + z))] = b
    return out


@pytest.mark.parametrize("bc,E,r,L", [("pbc", 1.0e5, 6, 16), ("free_space", 1.0e4, 5, 16), ("pbc", 1.0e5, 6, 15)])
def test_blocked_sweep_touches_only_reached_orders(bc, E, r, L):
    from coulomb_kmc_amd import KMCFMM
    rng = np.random.default_rng(71)
    N = 4000
    ncps = 2 ** (r - 1)
    pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
    q = np.ones(N)
    q[: N // 2] = -1.0
    rng.shuffle(q)
    kmcs = []
    for block in (1, 0):
        k = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=1.1 * E / ncps,
                   precision="fp64")   # (max_move: 1.1 leaf widths, hops below it)
        k.initialise()
        k.set_tiles(near_lanes=4, sweep_block=block, sweep_pairs=2, sweep_split=1)
        kmcs.append(k)
    blocked, fused = kmcs
    fused.set_expansions(blocked.get_expansions())   # the same start, bit for bit
    cuts_seen = set()
    for hop in range(3):
        old = pos[rng.integers(N)]
        new = old + rng.uniform(-0.6, 0.6, 3) * E / ncps
        new = np.where(new < -0.5 * E, new + E, np.where(new >= 0.5 * E, new - E, new)) if bc == "pbc" \
            else np.clip(new, -0.5 * E, 0.5 * E - 1e-9)
        before = blocked.get_expansions()
        blocked.sweep_hop(1.0, old, new)
        fused.sweep_hop(1.0, old, new)
        after, ref = blocked.get_expansions(), fused.get_expansions()
        cut = block_orders(E, ncps, L, bc == "pbc", old, new)
        cuts_seen.update(np.unique(cut).tolist())
        for c in np.unique(cut):
            leaves = np.nonzero(cut == c)[0]
            hi = np.r_[c * c: L * L, L * L + c * c: 2 * L * L]   # orders >= c, real then imaginary
            assert np.array_equal(after[np.ix_(leaves, hi)], before[np.ix_(leaves, hi)]), (bc, hop, c)
        scale = np.abs(ref).max()
        err = np.abs(after - ref).max()
        assert err <= 1e-13 * scale, (bc, hop, err, scale)
    assert min(cuts_seen) < L, cuts_seen   # some blocks really skip orders
    for k in kmcs:
        k.free()
