"""
Lattice bookkeeping on the device: Algorithm 2 of the reference's examples
(docs/source/example_bookkeeping.rst; examples/propose_with_dats_example.py:217-450),
with every part of a KMC step on the GPU:

1. proposals: each charge proposes its position plus M fixed offsets, wrapped into the
   domain (the example's ``prop_pos`` ParticleLoop);
2. overlap masking: a proposal within squared distance ``tol`` of another charge is
   masked off (the example's ``exclude`` PairLoop, TOL = 0.01);
3. the energy differences of all N*M proposals (``propose_with_dats``);
4. rates exp(-dE/kT) (the example's ``rate`` kernel; ``rate="metropolis"`` uses
   min(1, exp(-dE/kT))), 0 where masked, and the selection of one proposal from the
   running sum in (charge, offset) order against a uniform (``find_charge_to_move``);
5. the accept.

One call of ``kmc_lattice_step`` does 1-5 without host synchronisation, so
``capture`` can record many steps into one HIP graph.
"""

import ctypes
from itertools import product

import numpy as np

from . import _native

RATE_RULES = {"metropolis": 0, "exp": 1}


def example_offsets(la):
    """The example's M = 14 offsets (examples/propose_with_dats_example.py:55-84): the six
    first-shell and eight outer-corner points of a (-2..2)^3 stencil, times la / 3."""
    def keep(ox):
        if ox == (0, 0, 0):
            return False
        if sum(abs(v) for v in ox) == 1:
            return True
        return all(abs(v) == 2 for v in ox)
    offs = [ox for ox in product((-2, -1, 0, 1, 2), repeat=3) if keep(ox)]
    return np.array(offs, dtype=np.float64) * (la / 3.0)


def example_lattice(Ns, la):
    """The example's system (examples/propose_with_dats_example.py:162-180): Ns^3 sites of
    spacing la in a cube of side la*Ns, charges (-1)^i in (x, y, z) product order."""
    E = la * Ns
    s = 0.5 * E / Ns
    axis = [px * la - 0.5 * E + s for px in range(Ns)]
    pos = np.array(list(product(axis, axis, axis)), dtype=np.float64)
    q = np.array([(-1.0) ** i for i in range(len(pos))])
    return pos, q, E


class LatticeKMC:
    """Device-resident lattice KMC steps of one KMCFMM.

    :arg kmc: an initialised KMCFMM (its charge count is fixed for this object).
    :arg offsets: (M, 3) proposal offsets (components within max_move).
    :arg uniforms: uniforms in [0, 1), one per step, used cyclically.
    :arg tol: squared overlap distance (the example's TOL).
    :arg kT: rates exp(-dE/kT).
    :arg rate: "exp" (the example) or "metropolis".
    :arg wrap: wrap proposals into the domain (the example); False masks off
        proposals outside the domain.
    """

    def __init__(self, kmc, offsets, uniforms, tol=0.01, kT=1.0, rate="exp", wrap=True, precision=None):
        import torch
        if rate not in RATE_RULES:
            raise ValueError("rate must be 'exp' or 'metropolis'")
        kmc._assert_init()
        self.kmc = kmc
        self.offsets = np.ascontiguousarray(np.asarray(offsets, dtype=np.float64).reshape(-1, 3))
        self.M = self.offsets.shape[0]
        self.N = kmc.N
        self.T = self.N * self.M
        self.tol = float(tol)
        self.kT = float(kT)
        self.rate = RATE_RULES[rate]
        self.wrap = 1 if wrap else 0
        self.precision = precision or kmc.precision
        dev = "cuda"
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64).ravel())
        self.nsteps = u.size
        self.d_uniform = torch.as_tensor(u).to(dev)
        self.d_new_pos = torch.zeros(self.T * 3, dtype=torch.float64, device=dev)
        self.d_masks = torch.zeros(self.T, dtype=torch.int64, device=dev)
        self.d_dE = torch.zeros(self.T, dtype=torch.float64, device=dev)
        self.d_counter = torch.zeros(1, dtype=torch.int64, device=dev)
        self.d_selected = torch.full((self.nsteps,), -1, dtype=torch.int64, device=dev)
        self.graph = None

    def _call(self, stream):
        k = self.kmc
        if k.N != self.N:
            raise RuntimeError("the charge count changed since this LatticeKMC was made")
        _native.check(k._lib.kmc_lattice_step(
            k._ctx, stream, self.M, self.offsets.ctypes.data_as(ctypes.c_void_p), self.tol, self.wrap,
            self.d_new_pos.data_ptr(), self.d_masks.data_ptr(), self.d_dE.data_ptr(), self.nsteps,
            self.d_uniform.data_ptr(), self.kT, self.rate, self.d_counter.data_ptr(), self.d_selected.data_ptr(),
            _native.PREC[self.precision]), "kmc_lattice_step")

    def step(self, n=1):
        """n steps on torch's current stream (the first sizes the work buffers)."""
        import torch
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for _ in range(n):
            self._call(st)

    def capture(self, steps_per_graph=10):
        """Records steps_per_graph steps into one graph (after one eager step)."""
        import torch
        self.step(1)
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(steps_per_graph):
                self._call(ctypes.c_void_p(s.cuda_stream))
        self.graph = g
        return g

    def replay(self, n=1):
        for _ in range(n):
            self.graph.replay()

    def selected(self):
        """Proposal index i*M + m selected at each step so far (cyclic table)."""
        return self.d_selected.cpu().numpy()

    def steps_done(self):
        return int(self.d_counter.cpu()[0])

    def sync_host(self):
        """Copies the device energy and positions back into kmc.energy / kmc.positions;
        returns the device flags (4: a step found no positive rate, 2: a full leaf)."""
        k = self.kmc
        k.energy = k.device_energy()
        pos, _ = k.get_device_particles()
        k._pos()[:k.N] = pos
        return k.check_flags()
