"""
Tile autotuner for the propose/accept kernels.

Tunable shapes (set on a context with KMCFMM.set_tiles / kmc_set_tiles):
  near_lanes   lanes per evaluation point in the near-field kernel (1, 4, 8, 16, 32, 64;
               27 = one lane per neighbour leaf; 0 = a moving charge's points together;
               -G = points sorted by 4x4x4 leaf tile, each tile's charges staged in LDS
               once, G lanes per point)
  near_order   0: the per-point near field in the proposals' order; 1: over the points
               counting-sorted by 4x4x4 leaf tile, one contiguous range per XCD (same
               sums bit for bit)
  sweep_pairs  column pairs (m, L-1-m) per lane of the accept sweep (1, 2)
  sweep_block  leaves (threads) per workgroup of the accept sweep (64, 128, 192, 256), or
               0: fused layout, 64 leaves x all column groups per workgroup, with the
               expansion update staged through LDS and coalesced, or 1: blocked layout,
               a 4x4x4 block of leaves per workgroup with the orders of each source image
               truncated where they cannot change the block's potential by 2^-60
  sweep_split  waves sharing the 27 images of one leaf (1; or 3, one image layer per
               wave, workgroups of 64 leaves x 3; '27'/PBC only)

``autotune(kmc)`` times every candidate with HIP events (kmc_time_kernel) on the
context's current state and returns the fastest; ``choose_tiles`` reads the table of
tuned choices kept in tiles.json (keyed by expansion order, leaf count and proposal
batch size), falling back to defaults.  The sweep candidates are timed last because
timing the sweep re-applies the last accepted move.
"""

import json
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(_HERE, "tiles.json")

NEAR_LANES = (0, 1, 4, 8, 16, 27, 32, 64, -4, -8)
NEAR_SORTED_LANES = (4, 8, 16)   # candidates with near_order 1
# (sweep_pairs, sweep_block, sweep_split)
SWEEP_SHAPES = ((2, 1, 1), (1, 0, 1), (2, 0, 1), (1, 64, 1), (1, 256, 1), (2, 64, 1), (2, 256, 1), (1, 64, 3),
                (2, 64, 3))
DEFAULT = {"near_lanes": 27, "sweep_block": 0, "sweep_pairs": 2, "sweep_split": 1, "near_order": 0}
KEYS = ("near_lanes", "sweep_block", "sweep_pairs", "sweep_split", "near_order")


def _pow2(x):
    b = 1
    while b < x:
        b *= 2
    return b


def _key(L, T, ncells):
    # proposal batches and leaf counts are bucketed by powers of two
    return "L{}_C{}_T{}".format(int(L), _pow2(ncells), _pow2(T))


def load_table():
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            return json.load(f)
    return {}


def choose_tiles(L, T, ncells, **override):
    t = dict(DEFAULT)
    t.update(load_table().get(_key(L, T, ncells), {}))
    for k in KEYS:
        if override.get(k) is not None:
            t[k] = int(override[k])
    return {k: int(t[k]) for k in KEYS}


def sweep_candidates(periodic):
    return [s for s in SWEEP_SHAPES if periodic or s[2] == 1]


def autotune(kmc, T, reps=10, precision="mixed", save=False, log=None):
    """Time every candidate on kmc (a propose must have run); returns (choice, timings)."""
    timings = {"near": {}, "sweep": {}}
    cur = choose_tiles(kmc.L, T, kmc.ncells)
    cands = [(g, 0) for g in NEAR_LANES]
    if round(kmc.ncells ** (1.0 / 3.0)) % 4 == 0:
        cands += [(g, 1) for g in NEAR_SORTED_LANES]
    for g, o in cands:
        kmc.set_tiles(**dict(cur, near_lanes=g, near_order=o))
        kmc.time_kernel("near", reps=2, precision=precision)
        timings["near"]["{}/{}".format(g, o)] = kmc.time_kernel("near", reps=reps, precision=precision)
    best_g, best_o = (int(x) for x in min(timings["near"], key=timings["near"].get).split("/"))
    for npair, block, split in sweep_candidates(kmc.boundary_condition.value != "free_space"):
        kmc.set_tiles(near_lanes=best_g, sweep_block=block, sweep_pairs=npair, sweep_split=split, near_order=best_o)
        kmc.time_kernel("sweep", reps=1)
        timings["sweep"]["{}x{}x{}".format(npair, block, split)] = kmc.time_kernel("sweep", reps=reps)
    best = min(timings["sweep"], key=timings["sweep"].get)
    npair, block, split = (int(x) for x in best.split("x"))
    choice = {"near_lanes": int(best_g), "sweep_block": block, "sweep_pairs": npair, "sweep_split": split,
              "near_order": int(best_o)}
    kmc.set_tiles(**choice)
    if log:
        log("autotune", _key(kmc.L, T, kmc.ncells), choice, timings)
    if save:
        tab = load_table()
        tab[_key(kmc.L, T, kmc.ncells)] = choice
        with open(TABLE, "w") as f:
            json.dump(tab, f, indent=1, sort_keys=True)
    return choice, timings


def read_traffic(root, kernel, cfg):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled
    by a copy-kernel calibration), from profiles/traffic.json when present."""
    path = os.path.join(root, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tab = json.load(f)
    ent = tab.get("{}|{}".format(kernel, cfg["workload"]))
    return None if ent is None else ent.get("bytes_per_launch")
