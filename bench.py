#!/usr/bin/env python3
"""
FMM-KMC propose/accept throughput on one MI355X.

A step is one pass of the hot path over one batch of synthetic proposals: P charges
each propose k moves (displacement uniform in a ball, wrapped into the periodic box
or clipped to a free-space box), the energy differences of all T = P*k proposals
are computed, one proposal is selected by Metropolis (rates min(1, exp(-dE/kT)),
cumulative sum in proposal order, seeded uniform) and accepted (cell tables, every
leaf local expansion, the level-0 multipole).  Each step runs entirely on the
device; the steps are replayed from a captured HIP graph.

Prints one JSON line.  "value" is proposals/s of whole steps of the headline path:
the mixed-precision path when both paths hold their dE bars (FP64 1e-10, mixed 1e-8,
relative with |dE| floored at 1e-3) over all proposals of step 0 against the oracle's
extended-precision reference on the same state, else the FP64 path ("accuracy" says
which and why).  Accepts/s, both paths, their errors, the roofline of the dominant
kernel and the host-CPU baseline are in the same line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pbc-2m|pbc-100k|free-1k]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ΔE proposals/s + accepts/s on 1 GPU, mixed vs FP64, at ΔE rel err ≤1e-8"

# workloads (r = levels, leaf grid 2^(r-1) per side); steps/warmup are the defaults
WORKLOADS = {
    "free-1k": dict(workload="free-space N=1000 box 10 leaves 8^3 L=12, 64x4 proposals", N=1000, E=10.0, r=4,
                    L=12, bc="free_space", P=64, k=4, radius=1.0, seed=1000, steps=100, warmup=10, pipeline=1),
    "pbc-100k": dict(workload="pbc N=100000 box 100 leaves 32^3 L=12, 1024x8 proposals", N=100000, E=100.0, r=6,
                     L=12, bc="pbc", P=1024, k=8, radius=1.5, seed=1001, steps=1000, warmup=50, pipeline=2),
    "pbc-2m": dict(workload="pbc N=2000000 box 271.4 leaves 64^3 L=16, 32768x8 proposals", N=2000000, E=271.4,
                   r=7, L=16, bc="pbc", P=32768, k=8, radius=1.5, seed=1004, steps=100, warmup=20, pipeline=4),
}
DEFAULT_WORKLOAD = "pbc-2m"

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def sweep_flops_per_source(L):
    """FP64 flops of one P2L of one source into one leaf in k_sweep: 9 per packed
    coefficient (recurrence 5, complex accumulate 4), the power chain 6 per order,
    30 for displacement, 1/r and set-up."""
    return 9 * L * (L + 1) // 2 + 6 * (L - 1) + 30


SWEEP_TAU = 2.0 ** -60


SWEEP_BUCKETS = (8, 9, 10, 12)


def sweep_bucket_of(lc, L):
    """Orders the blocked sweep evaluates for an image of truncation order lc (its
    compile-time bucket; 0 for lc = 0)."""
    out = np.full(np.shape(lc), L)
    for b in reversed(SWEEP_BUCKETS):
        if b < L:
            out = np.where(lc <= b, b, out)
    return np.where(np.asarray(lc) <= 0, 0, out)


def sweep_work(E, ncps, L, periodic, old, new, blocked):
    """(FP64 flops, HBM bytes) of the accept sweep for the hop old -> new.  Flops: over
    every (leaf, source image) pair that is well separated (the others add nothing),
    sweep_flops_per_source of the image's order: L, or with the blocked layout
    (sweep_block 1) the truncation order lc of the image against the leaf's 4x4x4 block
    (the smallest lc with (rho/D)^lc / (D - rho) <= 2^-60, D the distance from the image
    to the block's leaf centres, rho a leaf's half diagonal), the rule the kernel
    applies.  Bytes: read + write of each leaf's packed expansion, all L(L+1)/2 entries,
    or with the blocked layout the entries of the orders below the bucket of the block's
    largest lc (the others receive no term and are not touched)."""
    w = E / ncps
    c0 = -0.5 * E + 0.5 * w
    rho = 0.5 * np.sqrt(3.0) * w
    ar = np.arange(ncps)
    tx, ty, tz = np.meshgrid(ar, ar, ar, indexing="ij")
    R = 1 if periodic else 0
    total = 0.0
    lcmax = np.zeros(tx.shape, int)
    for src in (old, new):
        src = np.asarray(src, dtype=np.float64)
        sc = np.minimum(((src + 0.5 * E) / w).astype(int), ncps - 1)
        for oz in range(-R, R + 1):
            for oy in range(-R, R + 1):
                for ox in range(-R, R + 1):
                    near = ((np.abs(sc[0] + ox * ncps - tx) < 2) & (np.abs(sc[1] + oy * ncps - ty) < 2) &
                            (np.abs(sc[2] + oz * ncps - tz) < 2))
                    if blocked:
                        s = src + np.array([ox, oy, oz]) * E
                        e = [np.maximum(np.abs(s[d] - (c0 + (4 * (t // 4) + 1.5) * w)) - 1.5 * w, 0.0)
                             for d, t in enumerate((tx, ty, tz))]
                        D = np.sqrt(e[0] ** 2 + e[1] ** 2 + e[2] ** 2)
                        far = D > 3 * rho
                        Ds = np.where(far, D, 1.0)
                        bound = np.where(far, 1.0 / np.where(far, Ds - rho, 1.0), np.inf)
                        lc = np.zeros(D.shape, int)
                        for _ in range(L):
                            more = (bound > SWEEP_TAU) & (lc < L)
                            bound = np.where(more, bound * (rho / Ds), bound)
                            lc = lc + more
                        lc = np.where(far, lc, L)
                    else:
                        lc = np.full(tx.shape, L)
                    lcmax = np.maximum(lcmax, lc)
                    f = 9 * lc * (lc + 1) // 2 + 6 * (lc - 1) + 30
                    total += float(np.sum(np.where(near, 0, f)))
    nb = sweep_bucket_of(lcmax, L) if blocked else np.full(tx.shape, L)
    return total, float(np.sum(2 * 16 * (nb * (nb + 1) // 2)))


def distinct_leaves(pts, E, ncps, periodic):
    """(leaves holding a point, leaves in the union of the points' 27-leaf stencils)."""
    w = E / ncps
    c = np.clip(((np.asarray(pts) + 0.5 * E) / w).astype(np.int64), 0, ncps - 1)
    own = np.unique(c[:, 0] + ncps * (c[:, 1] + ncps * c[:, 2]))
    seen = np.zeros(ncps ** 3, dtype=bool)
    ox = own % ncps
    oy = (own // ncps) % ncps
    oz = own // (ncps * ncps)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                x, y, z = ox + dx, oy + dy, oz + dz
                if periodic:
                    x, y, z = x % ncps, y % ncps, z % ncps
                    ok = np.ones(len(x), dtype=bool)
                else:
                    ok = (x >= 0) & (x < ncps) & (y >= 0) & (y < ncps) & (z >= 0) & (z < ncps)
                seen[(x + ncps * (y + ncps * z))[ok]] = True
    return len(own), int(seen.sum())


def kernel_rooflines(kt, N, ncells, L, P, T, bc, sweep_flops, sweep_bytes, leaves=None):
    """Per-kernel achieved rate against the attainable one, min(FP64 peak, AI x HBM peak),
    from the algorithmic work per launch (add/mul 1, FMA 2, rsqrt 1):
    near: 10 flops per pair, 27 rho pairs per evaluation point (T + P points, rho = N /
    leaves); bytes 32 per point in + 8 out, and 32 per charge of the leaves the points'
    stencils cover (leaves[1], counted on the host from a step's points);
    far: 9 nh flops per point; bytes 40 per point and 16 nh per distinct leaf holding a
    point (leaves[0]);
    sweep: as in the roofline line.  The long-range kernels read their operator through
    the scalar cache and wait on it and on memory latency, not on HBM bandwidth or the
    FP64 rate; they are reported as durations only."""
    nh = L * (L + 1) // 2
    rho = N / ncells
    ne = T + P
    work = {
        "near": (ne * 10.0 * 27 * rho, ne * 40.0 + (leaves[1] if leaves else min(ncells, 27 * P)) * rho * 32.0),
        "far": (ne * 9.0 * nh, ne * 40.0 + (leaves[0] if leaves else min(ncells, P)) * 16.0 * nh),
        "sweep": (float(sweep_flops), float(sweep_bytes)),
    }
    if bc == "ff-only":
        work.pop("near"), work.pop("far")
    out = {}
    for name, (fl, by) in work.items():
        ms = kt.get(name)
        if not ms:
            continue
        ai = fl / by
        att = min(FP64_PEAK_TFLOPS, ai * HBM_PEAK_GBS / 1e3)
        ach = fl / (ms * 1e-3) / 1e12
        out[name] = {"launch_ms": ms, "flops": fl, "bytes": by, "flop_per_byte": ai,
                     "bound": "fp64" if att >= FP64_PEAK_TFLOPS else "hbm",
                     "attainable_TFLOPs": att, "achieved_TFLOPs": ach, "frac": ach / att,
                     "hbm_GBs": by / (ms * 1e-3) / 1e9}
    return out


def make_system(cfg):
    rng = np.random.default_rng(cfg["seed"])
    N, E = cfg["N"], cfg["E"]
    pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
    q = np.ones(N)
    q[: N // 2] = -1.0
    rng.shuffle(q)
    return pos, q, rng


def move_stream(rng, cfg, nsteps):
    N, P, k, rad = cfg["N"], cfg["P"], cfg["k"], cfg["radius"]
    T = P * k
    pids = np.empty((nsteps, P), np.int64)
    for s in range(nsteps):
        pids[s] = rng.choice(N, size=P, replace=False)
    d = rng.normal(size=(nsteps, T, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    d *= rad * rng.uniform(size=(nsteps, T, 1)) ** (1.0 / 3.0)
    u = rng.uniform(size=nsteps)
    return pids, d, u


def host_moves(pos, pids_s, disp_s, k, E, bc):
    moves = []
    for j, pid in enumerate(pids_s):
        npos = pos[pid] + disp_s[j * k:(j + 1) * k]
        if bc == "free_space":
            npos = np.clip(npos, -0.5 * E, 0.5 * E)
        else:
            npos = np.where(npos < -0.5 * E, npos + E, npos)
            npos = np.where(npos >= 0.5 * E, npos - E, npos)
        moves.append((int(pid), npos))
    return moves


def time_propose_with_dats(kmc, pids_s, disp_s, k, E, bc, reps=5):
    """kmc_propose_with_dats at this workload's size, on device arrays [N][k] (the dats
    form of one step's proposals: the moving charges' masks set, the others clear);
    wall time per call (the call synchronises for its bad-move check) and the same
    proposals through kmc_propose for comparison."""
    import ctypes
    import torch
    from coulomb_kmc_amd import _native
    N = kmc.N
    pos, _ = kmc.get_device_particles()
    moves = host_moves(pos, pids_s, disp_s, k, E, bc)
    ppos = np.zeros((N, k, 3))
    masks = np.zeros((N, k), dtype=np.int64)
    for pid, npos in moves:
        ppos[pid] = npos
        masks[pid] = 1
    dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
    d_smc = dev(np.array([k], dtype=np.int64), torch.int64)
    d_sites = torch.zeros(N, dtype=torch.int64, device="cuda")
    d_ppos, d_masks = dev(ppos.reshape(-1), torch.float64), dev(masks.reshape(-1), torch.int64)
    d_out = torch.zeros(N * k, dtype=torch.float64, device="cuda")
    lib = _native.lib()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    tm = ctypes.c_int64(0)

    def call():
        _native.check(lib.kmc_propose_with_dats(kmc.handle, st, k, d_smc.data_ptr(), 1, d_sites.data_ptr(),
                                                d_ppos.data_ptr(), d_masks.data_ptr(), None, d_out.data_ptr(),
                                                _native.PREC[kmc.precision], ctypes.byref(tm)),
                      "kmc_propose_with_dats")
    call()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    ms = (time.time() - t0) / reps * 1e3
    # the same proposals through kmc_propose (device arrays, no dats set-up)
    P = len(pids_s)
    d_pid = dev(np.asarray(pids_s, dtype=np.int64), torch.int64)
    d_es = dev(np.arange(P + 1, dtype=np.int64) * k, torch.int64)
    d_np = dev(np.concatenate([m[1] for m in moves]).reshape(-1), torch.float64)
    d_dE = torch.zeros(P * k, dtype=torch.float64, device="cuda")
    kmc.propose_device(d_pid, d_es, d_np, d_dE)
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(reps):
        kmc.propose_device(d_pid, d_es, d_np, d_dE)
    torch.cuda.synchronize()
    ms_p = (time.time() - t0) / reps * 1e3
    same = bool(np.array_equal(d_out.view(N, k)[d_pid].cpu().numpy().reshape(-1), d_dE.cpu().numpy()))
    return {"ms_per_call": ms, "proposals": int(tm.value), "charges": N, "max_prop": k,
            "propose_ms_per_call": ms_p, "equal_to_propose": same,
            "note": "wall time per call with the device synchronisation it makes; dats arrays [N][max_prop] "
                    "on the device, the moving charges' masks set"}


def metropolis_pick(dE, u, kT=1.0):
    w = np.where(np.isfinite(dE), np.minimum(1.0, np.exp(-dE / kT)), 0.0)
    c = np.cumsum(w)
    return min(int(np.searchsorted(c, u * c[-1], side="right")), len(dE) - 1)


COPY_PROBE_BYTES = 512 * 2 ** 20


def measured_peaks(nbytes=2 ** 30, reps=10):
    """The roofline denominators measured on this device (kmc_probe_peaks): a streaming
    copy of nbytes (GB/s of read + write) and a loop of independent FP64 FMA chains."""
    import ctypes
    import torch
    from coulomb_kmc_amd import _native
    out = (ctypes.c_double * 2)()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _native.check(_native.lib().kmc_probe_peaks(st, nbytes, reps, out), "kmc_probe_peaks")
    return {"copy_GBs": out[0], "copy_bytes": nbytes, "fp64_fma_TFLOPs": out[1]}


def probe_hop(cfg):
    """The hop of a unit charge whose accept sweep the traffic probe measures (seeded)."""
    E = cfg["E"]
    rng = np.random.default_rng(cfg["seed"] + 11)
    old = rng.uniform(-0.5 * E, 0.5 * E, 3)
    new = old + np.array([0.6, -0.5, 0.4]) * cfg["radius"]
    if cfg["bc"] == "free_space":
        return old, np.clip(new, -0.5 * E, 0.5 * E)
    return old, np.where(new < -0.5 * E, new + E, np.where(new >= 0.5 * E, new - E, new))


def probe_traffic(kmc, cfg, pids_all, disp_all):
    """Launches for the HBM-traffic counters: one propose (sizes the evaluation arrays),
    5 accept sweeps, then 5 device copies of COPY_PROBE_BYTES (calibration of the
    counters against a kernel whose traffic is known: read + write of that size)."""
    import torch
    P, k = cfg["P"], cfg["k"]
    T = P * k
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    t_pids = torch.as_tensor(pids_all, dtype=torch.int64).cuda()
    t_disp = torch.as_tensor(disp_all, dtype=torch.float64).cuda()
    d_pid = torch.zeros(P, dtype=torch.int64, device="cuda")
    d_es = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
    d_np = torch.zeros(T * 3, dtype=torch.float64, device="cuda")
    d_dE = torch.zeros(T, dtype=torch.float64, device="cuda")
    kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
    kmc.propose_device(d_pid, d_es, d_np, d_dE, precision="mixed")    # sizes the work buffers
    kmc.step_device(d_pid, d_es, d_np, d_dE, torch.full((1,), 0.5, dtype=torch.float64, device="cuda"),
                    1.0, counter, torch.zeros(1, dtype=torch.int64, device="cuda"), precision="mixed")
    old, new = probe_hop(cfg)
    kmc.sweep_hop(1.0, old, new)   # the move the timed sweeps re-apply
    torch.cuda.synchronize()
    ms = kmc.time_kernel("sweep", reps=5)
    a = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log("probe: sweep", ms, "ms per launch; copy probe", COPY_PROBE_BYTES, "bytes x 5")


def _pmc_rows(d, counter):
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((r["Kernel_Name"], r.get("Dispatch_Id"), float(r["Counter_Value"])))
    if not rows:
        raise SystemExit("no {} rows under {}".format(counter, d))
    return rows


def _per_dispatch(rows):
    acc = {}
    for name, disp, v in rows:   # one row per dispatch (values summed over instances)
        acc.setdefault((name, disp), 0.0)
        acc[(name, disp)] += v
    return acc


def traffic_from_pmc(workload, fetch_dir, write_dir):
    """HBM bytes per accept-sweep launch from FETCH_SIZE / WRITE_SIZE (KiB per dispatch),
    each scaled by bytes / counter of the copy probe (which reads and writes
    COPY_PROBE_BYTES)."""
    cfg = WORKLOADS[workload]
    res = {}
    for which, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        per = _per_dispatch(_pmc_rows(d, counter))
        # the last 5 sweeps are the timed accept sweeps (initialise also launches k_sweep)
        sweep_rows = [(n, v) for (n, i), v in sorted(per.items(), key=lambda kv: int(kv[0][1] or 0))
                      if n.startswith("void k_sweep")][-5:]
        sweep = [v for _, v in sweep_rows]
        kname = sweep_rows[-1][0][len("void "):].split("<")[0] if sweep_rows else "k_sweep"
        big = sorted(((v, n) for (n, _), v in per.items() if not n.startswith("void k_")), reverse=True)
        copies = [v for v, n in big if v * 1024 > 0.5 * COPY_PROBE_BYTES]
        if not sweep or not copies:
            raise SystemExit("missing k_sweep or copy probe dispatches under " + d)
        kib = float(np.mean(sweep))
        cal = COPY_PROBE_BYTES / (float(np.median(copies)) * 1024.0)
        res[which] = {"counter": counter, "kernel": sweep_rows[-1][0], "KiB_per_dispatch": kib,
                      "dispatches": len(sweep),
                      "copy_probe_KiB": float(np.median(copies)), "scale": cal,
                      "bytes_per_launch": kib * 1024.0 * cal}
    path = os.path.join(ROOT, "profiles", "traffic.json")
    tab = {}
    if os.path.exists(path):
        with open(path) as f:
            tab = json.load(f)
    key = kname + "|" + cfg["workload"]
    old, new = probe_hop(cfg)
    _, alg = sweep_work(cfg["E"], 2 ** (cfg["r"] - 1), cfg["L"], cfg["bc"] != "free_space", old, new,
                        kname == "k_sweep_block")
    tab[key] = {"bytes_per_launch": res["fetch"]["bytes_per_launch"] + res["write"]["bytes_per_launch"],
                "fetch": res["fetch"], "write": res["write"],
                "algorithmic_bytes": alg, "hop": [list(old), list(new)],
                "copy_probe_bytes": COPY_PROBE_BYTES}
    with open(path, "w") as f:
        json.dump(tab, f, indent=1, sort_keys=True)
    print(json.dumps({key: tab[key]}, indent=1))


BATCH = dict(workload="batch: 256 systems, N log-uniform [2000, 200000] at density 0.1, half free-space half pbc, "
                      "8-16 charges per leaf, L in {8, 12, 16}, 256x8 proposals + 1 accept per system per step",
             systems=256, nmin=2000, nmax=200000, density=0.1, P=256, k=8, radius=1.0, seed=1003, steps=20, warmup=4,
             near_lanes=8)


def batch_systems(spec):
    """The systems of the batched workload (seeded)."""
    rng = np.random.default_rng(spec["seed"])
    out = []
    for i in range(spec["systems"]):
        N = int(round(np.exp(rng.uniform(np.log(spec["nmin"]), np.log(spec["nmax"])))))
        N -= N % 2
        E = (N / spec["density"]) ** (1.0 / 3.0)
        ncps = 2 ** int(round(np.log2((N / 12.0) ** (1.0 / 3.0))))
        ncps = min(32, max(8, ncps))
        r = int(np.log2(ncps)) + 1
        L = int(rng.choice((8, 12, 16)))
        bc = "free_space" if i % 2 == 0 else "pbc"
        pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
        q = np.ones(N)
        q[: N // 2] = -1.0
        rng.shuffle(q)
        out.append(dict(N=N, E=E, r=r, L=L, bc=bc, pos=pos, q=q, P=spec["P"], k=spec["k"], radius=spec["radius"]))
    return out, rng


def run_batch(args):
    import torch
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.batch import BatchSystem, KMCBatch
    spec = BATCH
    steps = args.steps if args.steps > 0 else spec["steps"]
    warm = args.warmup if args.warmup >= 0 else spec["warmup"]
    G = max(1, min(args.graph_steps, 2))
    steps = max(G, (steps // G) * G)
    warm = max(G, (warm // G) * G)
    t0 = time.time()
    specs, rng = batch_systems(spec)
    log("batch", len(specs), "systems, charges", sum(x["N"] for x in specs), "generated", round(time.time() - t0, 2), "s")

    def make(x):
        k = KMCFMM(x["pos"].copy(), x["q"].reshape(-1, 1), x["E"], r=x["r"], l=x["L"], boundary_condition=x["bc"],
                   max_move=x["radius"], precision="mixed", init_method="fmm")
        k.initialise()
        k.set_tiles(near_lanes=near_lanes, sweep_block=max(args.sweep_block, 0), sweep_pairs=2, sweep_split=1)
        return k

    near_lanes = args.near_lanes or spec["near_lanes"]
    nst = 16
    systems = []
    t0 = time.time()
    for x in specs:
        kmc = make(x)
        pids, disp, us = move_stream(rng, dict(N=x["N"], P=x["P"], k=x["k"], radius=x["radius"]), nst)
        dev = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).cuda()
        bs = BatchSystem(kmc, x["P"], x["k"], dev(pids, torch.int64), dev(disp.reshape(-1), torch.float64),
                         dev(us, torch.float64))
        bs.spec = x
        systems.append(bs)
    torch.cuda.synchronize()
    t_init = time.time() - t0
    log("initialised", len(systems), "systems in", round(t_init, 1), "s")
    batch = KMCBatch(systems, kT=1.0, precision="mixed", streams=args.batch_streams, mode=args.batch_mode)
    tiles = {}
    t0 = time.time()
    if args.batch_mode == "graph":
        tiles = batch.autotune(lambda s: make(s.spec), reps=3, log=log)
    elif not args.no_tune:
        choice, tuned = batch.autotune_kernels(log=log)
        tiles = {"choice": choice, "ms_per_step": {"near_lanes {} sweep_block {}".format(*k): v
                                                   for k, v in tuned.items()}}
    log("autotune", round(time.time() - t0, 1), "s")
    batch.capture(G)
    torch.cuda.synchronize()
    for _ in range(warm // G):
        batch.replay()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps // G):
        batch.replay()
    ev1.record()
    ev1.synchronize()
    ms = ev0.elapsed_time(ev1)
    flags = [f for f in batch.check_flags() if f]
    T = batch.proposals_per_step

    # CPU baseline: the oracle on a stated subset of the systems (smallest, median and
    # largest charge count, and the largest pbc system at L=16), each from its GPU state,
    # whole steps (propose of its 2048 proposals + Metropolis pick + accept)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
        nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(nth)
        order = np.argsort([x["N"] for x in specs])
        l16 = [i for i in order if specs[i]["L"] == 16 and specs[i]["bc"] == "pbc"]
        subset = sorted(set([int(order[0]), int(order[len(order) // 2]), int(order[-1])] + ([int(l16[-1])] if l16 else [])))
        t_cpu, n_prop, n_acc = 0.0, 0, 0
        for i in subset:
            x, bs = specs[i], systems[i]
            pos_g, _ = bs.kmc.get_device_particles()
            orc = O.Oracle(pos_g, x["q"], x["E"], x["r"], x["L"], bc=x["bc"], max_move=x["radius"],
                           lr=lr_operator_reference_layout(x["L"], x["E"]) if x["bc"] == "pbc" else None)
            orc.init_from_expansions(bs.kmc.get_expansions(), bs.kmc.device_energy())
            mrng = np.random.default_rng(100 + i)
            pids, disp, us = move_stream(mrng, dict(N=x["N"], P=x["P"], k=x["k"], radius=x["radius"]), 2)
            for st in range(2):
                moves = host_moves(orc.positions(), pids[st], disp[st], x["k"], x["E"], x["bc"])
                es = np.arange(x["P"] + 1, dtype=np.int64) * x["k"]
                tc = time.time()
                dc = orc.propose_arrays(pids[st], es, np.concatenate([m[1] for m in moves]))
                kc = metropolis_pick(dc, us[st])
                orc.accept(moves[kc // x["k"]][0], moves[kc // x["k"]][1][kc % x["k"]])
                t_cpu += time.time() - tc
                n_prop += x["P"] * x["k"]
                n_acc += 1
            del orc
        cpu = {"value": n_prop / t_cpu, "unit": "proposals/s", "cores": nth, "kind": "port",
               "accepts_per_s": n_acc / t_cpu,
               "sample": "2 whole steps (2048 proposals + 1 accept) of each of {} systems ({}; charges {}), "
                         "{:.1f} s, oracle/kmc_oracle.cpp (FP64, OpenMP); the CPU steps systems one after "
                         "another".format(len(subset), subset, [specs[i]["N"] for i in subset], t_cpu)}
        log("cpu baseline", cpu)

    # roofline of the dominant kernel, the batch's accept sweep: every system's move
    # record set to one hop of a charge (kmc_sweep_hop), then the batch's sweep launches
    # (one per order / periodicity group) timed with HIP events (kmc_batch_time_sweep);
    # bytes and flops of each system's sweep counted as for one system (sweep_work)
    sweep_blocked = tiles.get("choice", {}).get("sweep_block", max(args.sweep_block, 0)) == 1
    # an (order, periodicity) group sweeps blocked when every system of it can
    grp_blocked = {}
    for x in specs:
        key = (x["L"], x["bc"] != "free_space")
        grp_blocked[key] = grp_blocked.get(key, True) and sweep_blocked and (2 ** (x["r"] - 1)) % 4 == 0
    b_sw, f_sw = 0.0, 0.0
    hop_rng = np.random.default_rng(spec["seed"] + 7)
    for bs, x in zip(systems, specs):
        ncps = 2 ** (x["r"] - 1)
        E = x["E"]
        old = hop_rng.uniform(-0.5 * E, 0.5 * E, 3)
        new = old + np.array([0.5, -0.4, 0.3])
        new = np.where(new < -0.5 * E, new + E, np.where(new >= 0.5 * E, new - E, new)) if x["bc"] == "pbc" \
            else np.clip(new, -0.5 * E, 0.5 * E)
        bs.kmc.sweep_hop(1.0, old, new)
        f, b = sweep_work(E, ncps, x["L"], x["bc"] != "free_space", old, new,
                          blocked=grp_blocked[(x["L"], x["bc"] != "free_space")])
        f_sw += f
        b_sw += b
    torch.cuda.synchronize()
    t_sw = batch.time_sweep(reps=10)
    sw_gbs = b_sw / (t_sw * 1e-3) / 1e9
    sw_tf = f_sw / (t_sw * 1e-3) / 1e12
    roof = {"bound": "hbm", "kernel": "batched accept sweep ({}, one launch per order / periodicity group)".format(
                "kb_sweep_block" if sweep_blocked else "kb_sweep_fused"),
            "achieved": sw_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sw_gbs / HBM_PEAK_GBS,
            "traffic": None, "launch_ms": t_sw, "algorithmic_bytes": b_sw, "algorithmic_flops": f_sw,
            "fp64_TFLOPs": sw_tf, "fp64_frac": sw_tf / FP64_PEAK_TFLOPS,
            "note": "launch_ms: all groups' sweep launches of one batch step, back to back on one stream "
                    "(kmc_batch_time_sweep, HIP events); bytes and flops as bench.sweep_work for one hop per "
                    "system (blocked groups: only the orders that receive terms are read and written)"}
    log("batch sweep roofline", roof)
    out = {
        "metric": METRIC,
        "value": T * steps / (ms * 1e-3),
        "unit": "proposals/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform charges +-1, ball-displacement proposals, seeded Metropolis uniforms)",
        "config": {"workload": spec["workload"], "systems": len(systems), "charges": int(sum(x["N"] for x in specs)),
                   "proposals_per_step": T, "graph_steps": G, "streams": args.batch_streams,
                   "mode": args.batch_mode,
                   "tiles": ({"L{}_C{}_T{}".format(*k): v for k, v in tiles.items()}
                             if args.batch_mode == "graph" else tiles) or
                            "batched kernels: near field {}, sweep_block as set".format(
                                "one lane per neighbour leaf" if near_lanes in (0, 27) else
                                "{} lanes per point".format(near_lanes))},
        "accepts_per_s": len(systems) * steps / (ms * 1e-3),
        "initialise_s": t_init,
        "flags": flags,
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (T * steps / (ms * 1e-3)) / cpu["value"] if cpu else None,
    }
    print(json.dumps(out), flush=True)
    batch.free()
    for s in systems:
        s.kmc.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="0: the workload's default")
    ap.add_argument("--warmup", type=int, default=-1, help="-1: the workload's default")
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS) + ["batch-256"])
    ap.add_argument("--batch-streams", type=int, default=6,
                    help="streams of the batched workload: per-group launches on streams - 1 side streams (6)")
    ap.add_argument("--batch-mode", default="kernels", choices=("kernels", "graph"),
                    help="batched kernels (one launch per kernel kind) or per-system steps in one graph")
    ap.add_argument("--graph-steps", type=int, default=10, help="steps captured per HIP graph")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="batch: keep the given tiles (no autotuning)")
    ap.add_argument("--near-lanes", type=int, default=0, help="0: autotuned")
    ap.add_argument("--near-order", type=int, default=-1,
                    help="near-field entry order (kmc_set_near_order; -1: autotuned): 0 proposals, 1 leaf tiles")
    ap.add_argument("--sweep-block", type=int, default=-1,
                    help="accept-sweep layout (kmc_set_tiles sweep_block; -1: autotuned): 0 fused, 1 blocked with "
                         "order truncation, 64..256 leaves per workgroup")
    ap.add_argument("--pipeline", default="auto", choices=("auto", "on", "off", "setup", "setup-far", "sweep-main"),
                    help="on: kmc_run_steps with each step's sweep beside the next step's near field and long "
                         "range; setup: beside its set-up and long range only; setup-far: as setup, with the next "
                         "far field behind the sweep on its stream; sweep-main: near field, combine, select and "
                         "sweep on the caller's stream, tail + next set-up + long range on one side stream, far field "
                         "on another; off: kmc_materialize_moves + "
                         "kmc_step per step; auto: the workload's choice (on for free-1k, setup for pbc-100k, "
                         "sweep-main for pbc-2m, where they measured fastest)")
    ap.add_argument("--probe-traffic", action="store_true",
                    help="only launch the accept sweep and a copy of known size (for rocprofv3 --pmc runs)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="read two rocprofv3 --pmc outputs (FETCH_SIZE, WRITE_SIZE) of --probe-traffic, "
                         "write profiles/traffic.json and exit (no GPU)")
    args = ap.parse_args()

    if args.traffic_from:
        traffic_from_pmc(args.workload, *args.traffic_from)
        return
    if args.workload == "batch-256":
        if args.gpus != 1:
            print("bench.py: this build measures one GPU only; --gpus {} is not supported".format(args.gpus),
                  file=sys.stderr)
            sys.exit(2)
        run_batch(args)
        return

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (one MI355X per box); --gpus {} is not supported"
              .format(args.gpus), file=sys.stderr)
        sys.exit(2)

    import torch
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd import tune

    cfg = WORKLOADS[args.workload]
    if args.steps <= 0:
        args.steps = cfg["steps"]
    if args.warmup < 0:
        args.warmup = cfg["warmup"]
    N, E, r, L, bc, P, k = cfg["N"], cfg["E"], cfg["r"], cfg["L"], cfg["bc"], cfg["P"], cfg["k"]
    T = P * k
    G = max(1, args.graph_steps)
    steps = max(G, (args.steps // G) * G)
    warm = max(G, (args.warmup // G) * G)
    nstream = warm + 2 * steps + 16
    log("workload", args.workload, cfg["workload"], "steps", steps, "warmup", warm)

    pos0, q, rng = make_system(cfg)
    t0 = time.time()
    pids_all, disp_all, us = move_stream(rng, cfg, nstream)
    log("move stream", round(time.time() - t0, 2), "s")

    kmc = KMCFMM(pos0.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=cfg["radius"])
    tiles = tune.choose_tiles(L, T, (2 ** (r - 1)) ** 3, near_lanes=args.near_lanes or None,
                              sweep_block=None if args.sweep_block < 0 else args.sweep_block,
                              near_order=None if args.near_order < 0 else args.near_order)
    kmc.set_tiles(**tiles)
    torch.cuda.synchronize()
    t0 = time.time()
    kmc.initialise()
    torch.cuda.synchronize()
    t_init = time.time() - t0
    e_init = kmc.energy
    log("initialise", round(t_init, 2), "s, energy", e_init)
    if args.probe_traffic:
        probe_traffic(kmc, cfg, pids_all, disp_all)
        return

    dev = lambda x, t: torch.as_tensor(np.ascontiguousarray(x), dtype=t).cuda()
    t_pids = dev(pids_all, torch.int64)
    t_disp = dev(disp_all, torch.float64)
    t_u = dev(us, torch.float64)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    sel = torch.full((nstream,), -1, dtype=torch.int64, device="cuda")
    d_pid = torch.zeros(P, dtype=torch.int64, device="cuda")
    d_es = torch.zeros(P + 1, dtype=torch.int64, device="cuda")
    d_np = torch.zeros(T * 3, dtype=torch.float64, device="cuda")
    d_dE = torch.zeros(T, dtype=torch.float64, device="cuda")
    d_dE2 = torch.zeros(T, dtype=torch.float64, device="cuda")

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O   # the checker's error measure (and, below, the CPU baseline)

    # ---- dE of mixed vs FP64 on the initial state, 65536 proposals (or fewer)
    nerr = max(1, min(65536 // T, nstream))
    va, vm = [], []
    for b in range(nerr):
        counter.fill_(nstream - 1 - b)   # the last rows of the move stream, 0 <= row < nstream
        kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
        kmc.propose_device(d_pid, d_es, d_np, d_dE, precision="fp64")
        kmc.propose_device(d_pid, d_es, d_np, d_dE2, precision="mixed")
        va.append(d_dE.cpu().numpy())
        vm.append(d_dE2.cpu().numpy())
    derr = O.de_error(np.concatenate(vm), np.concatenate(va))
    log("mixed vs fp64 dE error", derr)

    # ---- step 0's proposals on both GPU paths (checked against the oracle's extended
    # reference below)
    counter.fill_(0)
    kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
    kmc.propose_device(d_pid, d_es, d_np, d_dE, precision="fp64")
    kmc.propose_device(d_pid, d_es, d_np, d_dE2, precision="mixed")
    g64_step0, gmx_step0 = d_dE.cpu().numpy(), d_dE2.cpu().numpy()
    accuracy = {"measure": "|dE - dE_ref| / max(|dE_ref|, {:g}) (energy units; kT = 1)".format(O.DE_FLOOR),
                "bars": {"fp64": O.DE_BAR_FP64, "mixed": O.DE_BAR_MIXED}}

    # ---- CPU baseline (oracle, same state and move stream), bounded sample
    cpu = None
    cpu_parity = None
    if not args.no_cpu:
        from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
        nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(nth)
        orc = O.Oracle(pos0, q, E, r, L, bc=bc, max_move=cfg["radius"],
                       lr=lr_operator_reference_layout(L, E) if bc in ("pbc", "ff-only") else None)
        orc.init_from_expansions(kmc.get_expansions(), kmc.energy)
        # extended-precision reference of step 0's proposals on the same state
        moves0 = host_moves(orc.positions(), pids_all[0], disp_all[0], k, E, bc)
        es0 = np.arange(P + 1, dtype=np.int64) * k
        npos0 = np.concatenate([m[1] for m in moves0])
        tc = time.time()
        ext0 = orc.propose_check(pids_all[0], es0, npos0, extended=True)
        t_ext = time.time() - tc
        accuracy["reference"] = ("oracle checker in long double (64-bit mantissa) on the GPU's initial state, "
                                 "long range in difference form; all {} proposals of step 0, {:.1f} s"
                                 .format(T, t_ext))
        accuracy["fp64"] = O.de_error(g64_step0, ext0)
        accuracy["mixed"] = O.de_error(gmx_step0, ext0)
        accuracy["fp64_pass"] = accuracy["fp64"]["max_rel"] <= O.DE_BAR_FP64
        accuracy["mixed_pass"] = accuracy["mixed"]["max_rel"] <= O.DE_BAR_MIXED
        log("accuracy vs extended reference", accuracy)
        t_cpu, n_cpu, cmax, cabs, cmed = 0.0, 0, 0.0, 0.0, 0.0
        gpos = pos0.copy()
        picks = []
        while n_cpu < 1 or (t_cpu < args.cpu_seconds and n_cpu < min(50, nstream - 1)):
            s = n_cpu
            moves = host_moves(orc.positions(), pids_all[s], disp_all[s], k, E, bc)
            es = np.arange(P + 1, dtype=np.int64) * k
            npos = np.concatenate([m[1] for m in moves])
            tc = time.time()
            dc = orc.propose_arrays(pids_all[s], es, npos)
            kc = metropolis_pick(dc, us[s])
            pid, p_new = moves[kc // k][0], moves[kc // k][1][kc % k]
            orc.accept(pid, p_new)
            t_cpu += time.time() - tc
            picks.append(kc)
            if s == 0:
                counter.fill_(0)
                kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
                kmc.propose_device(d_pid, d_es, d_np, d_dE, precision="fp64")
                g = d_dE.cpu().numpy()
                ref_order = O.de_error(g, dc)
                cmax, cabs, cmed = ref_order["max_rel"], ref_order["max_abs"], ref_order["median_rel"]
                # the reference-ordered FP64 restatement against the extended reference
                accuracy["cpu_reference_order"] = O.de_error(dc, ext0)
            n_cpu += 1
        cpu = {"value": n_cpu * T / t_cpu, "unit": "proposals/s", "cores": nth, "kind": "port",
               "accepts_per_s": n_cpu / t_cpu,
               "sample": "{} steps ({} proposals + 1 accept each) of this workload from the same state and move "
                         "stream, {:.1f} s, oracle/kmc_oracle.cpp (FP64, OpenMP, reference data structures)"
                         .format(n_cpu, T, t_cpu)}
        # one thread, propose only, on the first moving charges of the next step of the stream
        P1 = max(1, min(P, 2048))
        moves1 = host_moves(orc.positions(), pids_all[n_cpu][:P1], disp_all[n_cpu][:P1 * k], k, E, bc)
        O.set_threads(1)
        tc = time.time()
        orc.propose_arrays(pids_all[n_cpu][:P1], np.arange(P1 + 1, dtype=np.int64) * k,
                           np.concatenate([m[1] for m in moves1]))
        t1 = time.time() - tc
        O.set_threads(nth)
        cpu["single_thread"] = {"value": P1 * k / t1, "unit": "proposals/s",
                                "sample": "{} proposals ({} moving charges x {}) of the next step, propose only, "
                                          "{:.2f} s".format(P1 * k, P1, k, t1)}
        try:
            with open("/proc/cpuinfo") as fh:
                cpu["cpu_model"] = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")),
                                        "unknown")
        except OSError:
            cpu["cpu_model"] = "unknown"
        cpu_parity = {"gpu_fp64_vs_cpu_reference_order_step0": {"max_rel": cmax, "median_rel": cmed,
                                                                "max_abs": cabs, "rel_floor": O.DE_FLOOR},
                      "cpu_picks": picks}
        log("cpu baseline", cpu)

    # ---- timed runs: graph of G steps, replayed
    pipeline = {"auto": cfg["pipeline"], "on": 1, "off": 0, "setup": 2, "setup-far": 3, "sweep-main": 4}[args.pipeline]
    stream = torch.cuda.Stream()
    results = {}
    for prec in ("fp64", "mixed"):
        counter.fill_(0)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
            kmc.propose_device(d_pid, d_es, d_np, d_dE, precision=prec)   # sizes work buffers
        stream.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            if not pipeline:
                for _ in range(G):
                    kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
                    kmc.step_device(d_pid, d_es, d_np, d_dE, t_u, 1.0, counter, sel, precision=prec)
            else:
                # G steps in one call: each step's accept sweep beside the next step's
                # set-up, near field and long range
                kmc.run_steps(G, P, k, t_pids, t_disp, t_u, 1.0, counter, sel, d_pid, d_es, d_np, d_dE,
                              precision=prec, pipelined=int(pipeline), stream=stream.cuda_stream)
        counter.fill_(0)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            for _ in range(warm // G):
                graph.replay()
        stream.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        tw0 = time.time()
        with torch.cuda.stream(stream):
            ev0.record(stream)
            for _ in range(steps // G):
                graph.replay()
            ev1.record(stream)
        ev1.synchronize()
        wall = time.time() - tw0
        ms = ev0.elapsed_time(ev1)
        flags = kmc.check_flags()
        results[prec] = {"ms_per_step": ms / steps, "proposals_per_s": T * steps / (ms * 1e-3),
                         "accepts_per_s": steps / (ms * 1e-3), "wall_s": wall, "flags": flags,
                         "energy": kmc.device_energy()}
        log(prec, results[prec])
        del graph

    # ---- CPU/GPU accept-sequence agreement on the sampled steps (same state, same stream)
    if cpu_parity is not None:
        gpu_sel = []
        kchk = KMCFMM(pos0.copy(), q.reshape(-1, 1), E, r=r, l=L, boundary_condition=bc, max_move=cfg["radius"])
        kchk.set_tiles(**tiles)
        kchk.initialise()
        counter.fill_(0)
        sel.fill_(-1)
        for s in range(len(cpu_parity["cpu_picks"])):
            kchk.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
            if s == 0:
                kchk.propose_device(d_pid, d_es, d_np, d_dE, precision="fp64")
            kchk.step_device(d_pid, d_es, d_np, d_dE, t_u, 1.0, counter, sel, precision="fp64")
        torch.cuda.synchronize()
        gpu_sel = sel[:len(cpu_parity["cpu_picks"])].cpu().numpy().tolist()
        cpu_parity["gpu_picks"] = gpu_sel
        cpu_parity["accept_sequence_equal"] = gpu_sel == cpu_parity["cpu_picks"]
        kchk.free()
        log("accept sequence cpu == gpu:", cpu_parity["accept_sequence_equal"])

    # ---- dominant kernel: the accept sweep, timed alone with HIP events (changes state: last)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        counter.fill_(0)
        kmc.materialize_moves(P, k, t_pids, t_disp, counter, d_pid, d_es, d_np)
        kmc.propose_device(d_pid, d_es, d_np, d_dE, precision="mixed", stream=torch.cuda.current_stream().cuda_stream)
    stream.synchronize()
    kt = {}
    for name in ("near", "far", "lr"):
        kt[name] = kmc.time_kernel(name, reps=20, precision="mixed")
    # the accept sweep, timed on three hops of a charge (kmc_sweep_hop sets the move that
    # kmc_time_kernel re-applies); its flops counted for each hop by sweep_work
    ncps = 2 ** (r - 1)
    ncells = ncps ** 3
    blocked = tiles["sweep_block"] == 1 and ncps % 4 == 0 and bc != "ff-only"
    sweep_kernel = "k_sweep_block" if blocked else "k_sweep_fused"
    hop_rng = np.random.default_rng(cfg["seed"] + 7)
    t_sum, f_sum, b_sum = 0.0, 0.0, 0.0
    for _ in range(3):
        old = hop_rng.uniform(-0.5 * E, 0.5 * E, 3)
        d = hop_rng.normal(size=3)
        new = old + cfg["radius"] * d / np.linalg.norm(d) * hop_rng.uniform() ** (1.0 / 3.0)
        new = np.where(new < -0.5 * E, new + E, np.where(new >= 0.5 * E, new - E, new)) if bc != "free_space" \
            else np.clip(new, -0.5 * E, 0.5 * E)
        kmc.sweep_hop(1.0, old, new)
        t_sum += kmc.time_kernel("sweep", reps=10)
        f, b = sweep_work(E, ncps, L, bc != "free_space", old, new, blocked)
        f_sum += f
        b_sum += b
    kt["sweep"] = t_sum / 3
    sweep_flops = f_sum / 3
    log("kernel ms", kt)
    nimg = 54 if bc != "free_space" else 2
    sweep_bytes = b_sum / 3
    ach_tflops = sweep_flops / (kt["sweep"] * 1e-3) / 1e12
    ach_gbs = sweep_bytes / (kt["sweep"] * 1e-3) / 1e9
    ai = sweep_flops / sweep_bytes
    hbm_bound = ai * HBM_PEAK_GBS / 1e3 < FP64_PEAK_TFLOPS
    flop_note = ("flops = sum over (leaf, well-separated source image) of 9 lc(lc+1)/2 + 6(lc-1) + 30, lc the "
                 "image's truncation order against the leaf's 4x4x4 block (mean of 3 hops; all L orders would be "
                 "{:.3g})".format(ncells * nimg * sweep_flops_per_source(L)) if blocked else
                 "flops = leaves x {} images x (9 L(L+1)/2 + 6(L-1) + 30)".format(nimg))
    if hbm_bound:
        roof = {"bound": "hbm", "kernel": "accept sweep ({})".format(sweep_kernel),
                "achieved": ach_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach_gbs / HBM_PEAK_GBS,
                "fp64_TFLOPs": ach_tflops, "fp64_frac": ach_tflops / FP64_PEAK_TFLOPS}
    else:
        roof = {"bound": "valu", "kernel": "accept sweep ({})".format(sweep_kernel),
                "achieved": ach_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach_tflops / FP64_PEAK_TFLOPS, "hbm_GBs": ach_gbs, "hbm_frac": ach_gbs / HBM_PEAK_GBS}
    roof.update({"traffic": None, "launch_ms": kt["sweep"], "algorithmic_flops": sweep_flops,
                 "algorithmic_bytes": sweep_bytes, "flop_per_byte": ai,
                 "note": "no dense contraction: MFMA unused; bytes = read + write of every leaf's packed "
                         "expansion; " + flop_note})
    traffic = tune.read_traffic(ROOT, sweep_kernel, cfg)
    if traffic is not None:
        roof["traffic"] = traffic
    peaks = measured_peaks()
    log("measured peaks", peaks)
    roof["measured_peaks"] = peaks
    roof["frac_of_measured"] = (ach_gbs / peaks["copy_GBs"]) if hbm_bound else (ach_tflops / peaks["fp64_fma_TFLOPs"])

    dats = time_propose_with_dats(kmc, pids_all[0], disp_all[0], k, E, bc)
    log("propose_with_dats", dats)
    mixed, fp64 = results["mixed"], results["fp64"]
    pos_now, _ = kmc.get_device_particles()
    mv0 = host_moves(pos_now, pids_all[0], disp_all[0], k, E, bc)
    pts0 = np.concatenate([np.concatenate([m[1] for m in mv0]), pos_now[np.asarray(pids_all[0])]])
    leaves = distinct_leaves(pts0, E, 2 ** (r - 1), bc != "free_space")
    kroof = kernel_rooflines(kt, N, ncells, L, P, T, bc, sweep_flops, sweep_bytes, leaves)
    kroof["distinct_leaves"] = {"holding_a_point": leaves[0], "in_the_stencils": leaves[1],
                                "note": "one step's T + P evaluation points"}
    # the headline path: mixed only when both paths hold their bars against the extended
    # reference; otherwise the FP64 path (whose bar is the tighter one), or nothing
    # checked at all without the oracle (--no-cpu)
    checked = "mixed_pass" in accuracy
    if checked and accuracy["mixed_pass"] and accuracy["fp64_pass"]:
        head = "mixed"
    elif checked and accuracy["fp64_pass"]:
        head = "fp64"
    else:
        head = "unchecked" if not checked else "failed"
    accuracy["headline_path"] = head
    hv = mixed if head in ("mixed", "unchecked") else fp64
    out = {
        "metric": METRIC,
        "value": hv["proposals_per_s"] if head != "failed" else None,
        "unit": "proposals/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": hv["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform charges +-1, ball-displacement proposals, seeded Metropolis uniforms)",
        "config": {"workload": cfg["workload"], "N": N, "box": E, "leaves_per_side": 2 ** (r - 1), "r": r,
                   "L": L, "bc": bc, "proposals_per_step": T, "moving_charges_per_step": P,
                   "ball_radius": cfg["radius"], "kT": 1.0, "graph_steps": G, "tiles": tiles,
                   "pipelined": pipeline},
        "accepts_per_s": hv["accepts_per_s"],
        "accuracy": accuracy,
        "mixed": mixed,
        "fp64": fp64,
        "mixed_vs_fp64": mixed["proposals_per_s"] / fp64["proposals_per_s"],
        "dE_error_mixed_vs_fp64": derr,
        "kernel_ms": kt,
        "kernel_roofline": kroof,
        "propose_with_dats": dats,
        "initialise_s": t_init,
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (hv["proposals_per_s"] / cpu["value"]) if cpu else None,
        "parity": cpu_parity,
    }
    print(json.dumps(out), flush=True)
    kmc.free()


if __name__ == "__main__":
    main()
