"""
Many independent systems stepped together on one GPU.

Each system is its own KMCFMM context (its own leaf expansions, cell tables and work
buffers).  One step of the batch is one KMC step of every system: the synthetic moves
are materialised on the device, the proposals evaluated, one proposal selected by
Metropolis and accepted.  Two ways to run it:

* ``mode="kernels"`` (default): kmc_batch_step launches each kernel kind once for the
  whole batch (once per expansion order for the order-specialised kernels), the
  system being blockIdx.z of a table of per-system kernel arguments, so a batch step
  is ~20 launches whatever the number of systems.  The accept sweep of an (order,
  periodicity) group runs the blocked layout with order truncation over a flat table
  of (system, 4x4x4 leaf block) pairs when every system of the group is set to
  sweep_block 1 (no workgroup sized for a larger system), else the fused layout with
  two column pairs per lane; the near field runs G lanes per point when every system
  selects the same G with set_tiles(near_lanes=G), else one lane per neighbour leaf.
  ff-only systems form groups of their own (long range and cell tables only).
  ``autotune_kernels`` picks the common near-field lane count and sweep layout by
  timing whole batch steps.
* ``mode="graph"``: the systems' own kmc_step calls one after the other, with each
  system's tiles from the autotuner (per order / leaf count / proposal count class).

Either way ``capture`` records ``steps_per_graph`` batch steps into one HIP graph, a
batch step is one graph launch from the host.  In kernels mode ``streams=k`` sends
the per-group launches (far field, long range, sweep and cell-table tail of each
(order, periodicity) group) to k - 1 library-owned side streams, forked from and
joined to the step's stream with events inside kmc_batch_step, so the groups
overlap.  Graph mode records the systems on one stream: recording them on extra
torch side streams (forked from and joined to the capturing stream) ended in a host
crash inside graph instantiation on this ROCm/PyTorch stack, so that option raises.
"""

import ctypes

import numpy as np

from . import _native, tune


class BatchSystem:
    """One system of a batch: the context and its device move stream.

    pids_all [nsteps * P] int64, disp_all [nsteps * P * k * 3] float64, uniforms [nsteps]
    float64 are device tensors; the step counter selects row (counter mod nsteps)."""

    def __init__(self, kmc, P, k, pids_all, disp_all, uniforms):
        import torch
        self.kmc = kmc
        self.P, self.k = int(P), int(k)
        self.T = self.P * self.k
        self.pids_all, self.disp_all, self.uniforms = pids_all, disp_all, uniforms
        dev = pids_all.device
        self.counter = torch.zeros(1, dtype=torch.int64, device=dev)
        self.selected = torch.full((uniforms.numel(),), -1, dtype=torch.int64, device=dev)
        self.d_pid = torch.zeros(self.P, dtype=torch.int64, device=dev)
        self.d_es = torch.zeros(self.P + 1, dtype=torch.int64, device=dev)
        self.d_new = torch.zeros(self.T * 3, dtype=torch.float64, device=dev)
        self.d_dE = torch.zeros(self.T, dtype=torch.float64, device=dev)

    def materialize(self, stream=None):
        self.kmc.materialize_moves(self.P, self.k, self.pids_all, self.disp_all, self.counter, self.d_pid, self.d_es,
                                   self.d_new, stream=stream)

    def step(self, kT, precision, stream=None):
        self.materialize(stream)
        self.kmc.step_device(self.d_pid, self.d_es, self.d_new, self.d_dE, self.uniforms, kT, self.counter,
                             self.selected, precision=precision, stream=stream)


class KMCBatch:
    def __init__(self, systems, kT=1.0, precision="mixed", streams=1, mode="kernels"):
        if mode not in ("kernels", "graph"):
            raise ValueError("mode must be 'kernels' or 'graph'")
        if mode == "graph" and int(streams) != 1:
            raise NotImplementedError("graph-mode batches are recorded on one stream (recording the systems on "
                                      "extra torch side streams crashed graph instantiation on this ROCm/PyTorch "
                                      "stack)")
        if not 1 <= int(streams) <= 16:
            raise ValueError("streams must be in [1, 16]")
        self.streams = int(streams)
        self.systems = list(systems)
        self.kT = float(kT)
        self.precision = precision
        self.mode = mode
        self.graph = None
        self.steps_per_graph = 0
        self._batch = None
        if mode == "kernels" and self.systems:
            lib = _native.lib()
            n = len(self.systems)
            arr = (ctypes.c_void_p * n)(*[s.kmc.handle.value for s in self.systems])
            h = ctypes.c_void_p()
            _native.check(lib.kmc_batch_create(ctypes.byref(h), arr, n), "kmc_batch_create")
            self._batch = h
            if self.streams > 1:
                _native.check(lib.kmc_batch_set_streams(h, self.streams), "kmc_batch_set_streams")
            for i, s in enumerate(self.systems):
                _native.check(lib.kmc_batch_set_system(
                    h, i, s.uniforms.numel(), s.P, s.k, s.pids_all.data_ptr(), s.disp_all.data_ptr(),
                    s.uniforms.data_ptr(), s.counter.data_ptr(), s.selected.data_ptr(), s.d_pid.data_ptr(),
                    s.d_es.data_ptr(), s.d_new.data_ptr(), s.d_dE.data_ptr()), "kmc_batch_set_system")

    def free(self):
        if self._batch is not None and self._batch.value:
            _native.lib().kmc_batch_destroy(self._batch)
        self._batch = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def proposals_per_step(self):
        return sum(s.T for s in self.systems)

    def autotune(self, make_copy, reps=5, log=None):
        """Tiles per (L, leaves, proposals) class.  make_copy(system) returns a fresh,
        initialised KMCFMM of the same charges: timing the sweep re-applies a move, so
        the timings run on that copy and the batch's own systems keep their state."""
        import torch
        classes = {}
        for s in self.systems:
            classes.setdefault((s.kmc.L, s.kmc.ncells, s.T), []).append(s)
        choices = {}
        for key, members in classes.items():
            s0 = members[0]
            tmp = BatchSystem(make_copy(s0), s0.P, s0.k, s0.pids_all, s0.disp_all, s0.uniforms)
            tmp.materialize()
            tmp.kmc.propose_device(tmp.d_pid, tmp.d_es, tmp.d_new, tmp.d_dE, precision=self.precision)
            tmp.step(self.kT, self.precision)
            tmp.materialize()
            tmp.kmc.propose_device(tmp.d_pid, tmp.d_es, tmp.d_new, tmp.d_dE, precision=self.precision)
            torch.cuda.synchronize()
            choice, _ = tune.autotune(tmp.kmc, tmp.T, reps=reps, precision=self.precision)
            tmp.kmc.free()
            for s in members:
                s.kmc.set_tiles(**choice)
            choices[key] = choice
            if log:
                log("batch tiles", key, choice)
        return choices

    def autotune_kernels(self, near_lanes=(4, 8, 16), sweep_blocks=(0, 1), steps=4, log=None):
        """Kernels mode: times whole batch steps (captured, steps_per_graph 2) for every
        common near-field lane count and accept-sweep layout and keeps the fastest (the
        batched kernels take one near_lanes for all systems and one sweep layout per
        group).  The timed steps advance the systems like any other steps."""
        import torch
        if self.mode != "kernels":
            raise ValueError("autotune_kernels is for kernels mode (graph mode: autotune)")
        timings = {}
        for g in near_lanes:
            for sb in sweep_blocks:
                for s in self.systems:
                    s.kmc.set_tiles(near_lanes=g, sweep_block=sb, sweep_pairs=2, sweep_split=1)
                self.capture(2)
                self.replay(1)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                self.replay(max(1, steps // 2))
                e1.record()
                e1.synchronize()
                timings[(g, sb)] = e0.elapsed_time(e1) / (2 * max(1, steps // 2))
                if log:
                    log("batch tiles near_lanes {} sweep_block {}: {:.3f} ms/step".format(g, sb, timings[(g, sb)]))
        best = min(timings, key=timings.get)
        for s in self.systems:
            s.kmc.set_tiles(near_lanes=best[0], sweep_block=best[1], sweep_pairs=2, sweep_split=1)
        self.graph = None
        return {"near_lanes": best[0], "sweep_block": best[1]}, timings

    def prepare(self):
        """Sizes every context's work buffers (and, in kernels mode, uploads the argument
        tables); allocates, so it runs outside capture."""
        import torch
        if self.mode == "kernels":
            _native.check(_native.lib().kmc_batch_prepare(self._batch, self.kT, _native.PREC[self.precision]),
                          "kmc_batch_prepare")
        else:
            for s in self.systems:
                s.materialize()
                s.kmc.propose_device(s.d_pid, s.d_es, s.d_new, s.d_dE, precision=self.precision)
        torch.cuda.synchronize()

    def step(self, stream=None):
        """One KMC step of every system on the stream (torch's current stream by default)."""
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        if self.mode == "kernels":
            _native.check(_native.lib().kmc_batch_step(self._batch, ctypes.c_void_p(st)), "kmc_batch_step")
        else:
            for s in self.systems:
                s.step(self.kT, self.precision, stream=st)

    def capture(self, steps_per_graph=1):
        """Records steps_per_graph batch steps into one graph."""
        import torch
        self.prepare()
        main = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=main):
            for _ in range(steps_per_graph):
                self.step(main.cuda_stream)
        self.graph = g
        self.steps_per_graph = steps_per_graph
        return g

    def replay(self, n=1, reserve=False):
        """n replays of the captured steps.  reserve: before each replay every system gets
        room for its steps (KMCFMM.reserve_store, one synchronisation per replay) and the
        graph is captured again when a system's cell tables grew, so no selected move is
        ever dropped for a full leaf."""
        for _ in range(n):
            if reserve:
                self.reserve(self.steps_per_graph)
            self.graph.replay()

    def time_sweep(self, reps=5, stream=None):
        """Average ms of the batch's accept-sweep launches (kernels mode; re-applies every
        system's last accepted move, so the leaf expansions change)."""
        import torch
        if self._batch is None:
            raise ValueError("time_sweep is for kernels mode")
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        ms = ctypes.c_double(0.0)
        _native.check(_native.lib().kmc_batch_time_sweep(self._batch, ctypes.c_void_p(st), int(reps),
                                                         ctypes.byref(ms)), "kmc_batch_time_sweep")
        return ms.value

    def reserve(self, slots):
        """Room for `slots` more steps in every system's cell tables; captures the graph
        again (same steps per graph) when any grew.  Returns whether any grew."""
        import torch
        torch.cuda.synchronize()
        grew = False
        for s in self.systems:
            grew |= s.kmc.reserve_store(slots)
        if grew and self.graph is not None:
            self.capture(self.steps_per_graph)
        return grew

    def check_flags(self):
        return [s.kmc.check_flags() for s in self.systems]

    def energies(self):
        return np.array([s.kmc.device_energy() for s in self.systems])
