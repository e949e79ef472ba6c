"""
Tile autotuner for the propose/accept kernels.

Tunable shapes (set on a context with KMCFMM.set_tiles / kmc_set_tiles):
  near_lanes   lanes per evaluation point in the near-field kernel (1, 4, 8, 16, 32, 64;
               27 = one lane per neighbour leaf)
  sweep_block  threads per block of the accept sweep (64, 128, 192, 256)
  sweep_pairs  column pairs per lane of the accept sweep (1, 2; 3 for L <= 12)

``autotune(kmc)`` times every candidate with HIP events (kmc_time_kernel) on the
context's current state and returns the fastest; ``choose_tiles`` reads the table of
tuned choices kept in tiles.json (keyed by expansion order and proposal batch
size), falling back to defaults.  The sweep candidates are timed last because timing
the sweep re-applies the last accepted move.
"""

import json
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(_HERE, "tiles.json")

NEAR_LANES = (1, 4, 8, 16, 27, 32, 64)
SWEEP_BLOCKS = (64, 128, 256)
DEFAULT = {"near_lanes": 27, "sweep_block": 256, "sweep_pairs": 2}


def _key(L, T):
    # proposal batches are bucketed by powers of two
    b = 1
    while b < T:
        b *= 2
    return "L{}_T{}".format(int(L), b)


def load_table():
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            return json.load(f)
    return {}


def choose_tiles(L, T, near_lanes=None, sweep_block=None, sweep_pairs=None):
    t = dict(DEFAULT)
    t.update(load_table().get(_key(L, T), {}))
    if near_lanes:
        t["near_lanes"] = int(near_lanes)
    if sweep_block:
        t["sweep_block"] = int(sweep_block)
    if sweep_pairs:
        t["sweep_pairs"] = int(sweep_pairs)
    return {"near_lanes": int(t["near_lanes"]), "sweep_block": int(t["sweep_block"]),
            "sweep_pairs": int(t["sweep_pairs"])}


def autotune(kmc, T, reps=10, precision="mixed", save=False, log=None):
    """Time every candidate on kmc (a propose must have run); returns (choice, timings)."""
    timings = {"near": {}, "sweep": {}}
    cur = choose_tiles(kmc.L, T)
    for g in NEAR_LANES:
        kmc.set_tiles(near_lanes=g, sweep_block=cur["sweep_block"], sweep_pairs=cur["sweep_pairs"])
        kmc.time_kernel("near", reps=2, precision=precision)
        timings["near"][g] = kmc.time_kernel("near", reps=reps, precision=precision)
    best_g = min(timings["near"], key=timings["near"].get)
    for npair in ((1, 2, 3) if kmc.L <= 12 else (1, 2)):
        for b in SWEEP_BLOCKS:
            kmc.set_tiles(near_lanes=best_g, sweep_block=b, sweep_pairs=npair)
            kmc.time_kernel("sweep", reps=1)
            timings["sweep"]["{}x{}".format(npair, b)] = kmc.time_kernel("sweep", reps=reps)
    best = min(timings["sweep"], key=timings["sweep"].get)
    best_np, best_b = (int(x) for x in best.split("x"))
    choice = {"near_lanes": int(best_g), "sweep_block": best_b, "sweep_pairs": best_np}
    kmc.set_tiles(**choice)
    if log:
        log("autotune", _key(kmc.L, T), choice, timings)
    if save:
        tab = load_table()
        tab[_key(kmc.L, T)] = choice
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
