"""
KMCFMM: drop-in replacement of coulomb_kmc.KMCFMM (coulomb_kmc/kmc_fmm.py:408-1125)
for the FMM-KMC propose/accept path, backed by the HIP library (libkmc_amd.so).

Same constructor, same method names, argument meaning, return values and error types:

    kmc = KMCFMM(positions, charges, domain, r=..., l=..., boundary_condition='pbc')
    kmc.initialise()                 # energy in kmc.energy
    kmc.propose(((pid, new_positions), ...))          -> tuple of total energies
    kmc.propose_with_dats(site_max_counts, current_sites, prop_positions,
                          prop_masks, prop_energy_diffs)  # dE written in place
    kmc.accept((pid, new_position))  # moves the charge in `positions`
    kmc.free()

Arrays take the place of ppmd's ParticleDat/PositionDat/ScalarArray (only their
(npart, ncomp) contents are used on this path).  Extra keyword ``precision``
selects 'fp64' (default, all FP64) or 'mixed' for the propose kernels.
"""

import ctypes
import math
import time
from functools import lru_cache

import numpy as np

from . import _native
from .common import BCType, Domain, leaf_of
from . import lattice
from .kmc_inject_extract import InjectorExtractor

REAL = ctypes.c_double
INT64 = ctypes.c_int64


def _as_array(obj, ncomp, dtype):
    """(n, ncomp) numpy view of a dat-like object or array (no copy when possible)."""
    if isinstance(obj, np.ndarray):
        a = obj
    elif hasattr(obj, "view"):
        a = np.asarray(obj.view)
    elif hasattr(obj, "data") and not isinstance(obj.data, memoryview):
        a = np.asarray(obj.data)
    else:
        a = np.asarray(obj)
    if a.ndim == 1:
        a = a.reshape(-1, ncomp)
    return a


INIT_METHODS = {"auto": 0, "direct": 1, "fmm": 2}


@lru_cache(maxsize=16)
def _lr_operator_csr(L, extent):
    G = lattice.half_operator(L, extent)
    return lattice.csr(G)


@lru_cache(maxsize=16)
def lr_operator_reference_layout(L, extent):
    """(data, indptr, indices, dipole) of R in the reference layout (for the CPU restatement)."""
    R = lattice.far_field_matrix(L, extent)
    d, p, i = lattice.csr(R)
    return d, p, i, lattice.dipole_correction(extent)


def default_levels(N):
    """Levels r (leaf grid 2^(r-1)) giving at most ~16 charges per leaf, r >= 3."""
    r = 3
    while N / float(8 ** (r - 1)) > 16.0 and r < 11:
        r += 1
    return r


class KMCFMM(InjectorExtractor):
    """
    Electrostatic energy differences for kinetic Monte Carlo with the FMM-KMC
    algorithm, on one MI355X.

    :arg positions: (N, 3) float64 charge positions (updated in place by accept).
    :arg charges: (N, 1) or (N,) float64 charge values.
    :arg domain: object with ``extent`` (cubic box centred on the origin), or a float.
    :arg int N: number of charges (default: len(positions)).
    :arg str boundary_condition: 'pbc', 'free_space', '27' or 'ff-only'.
    :arg int r: number of FMM levels (leaf grid 2^(r-1) per side).
    :arg float energy_unit: multiplies every energy.
    :arg int l: number of expansion terms L (2..20).
    :arg float max_move: maximum distance of a proposed move (default: the extent).
    :arg str precision: 'fp64' or 'mixed' arithmetic for the propose kernels.
    :arg str init_method: how ``initialise`` builds the leaf expansions: 'direct' (exact
        P2L of every charge into every leaf), 'fmm' (hierarchical, M2L truncation at
        order l, needs r >= 4) or 'auto' (hierarchical for large systems).
    """

    def __init__(self, positions, charges, domain, N=None, boundary_condition="pbc", r=None, shell_width=0.0,
                 energy_unit=1.0, _debug=False, l=None, max_move=None, cuda_direct=False, mirror_direction=None,
                 precision="fp64", occ_headroom=8, init_method="auto"):
        if mirror_direction is not None:
            raise NotImplementedError("mirror charges (Dirichlet boundaries) are not part of this implementation")
        self._bc = BCType(boundary_condition)
        self.boundary_condition = self._bc
        if not hasattr(domain, "extent"):
            domain = Domain(domain)
        ext = np.asarray(domain.extent, dtype=np.float64)
        if not (abs(ext[0] - ext[1]) < 1e-12 * ext[0] and abs(ext[0] - ext[2]) < 1e-12 * ext[0]):
            raise AssertionError("only cubic domains are supported")
        self.domain = domain
        self.positions = positions
        self.charges = charges
        # the caller's objects (inject accepts them as keys after positions / charges
        # have been replaced by host arrays)
        self._user_positions = positions
        self._user_charges = charges
        pos = _as_array(positions, 3, np.float64)
        self._N_given = N is not None
        self.N = int(N) if N is not None else int(pos.shape[0])
        self.extent = float(ext[0])
        self.R = int(r) if r is not None else default_levels(self.N)
        self.L = int(l) if l is not None else 12
        self.energy_unit = float(energy_unit)
        self.max_move = float(max_move) if max_move is not None else float(max(ext))
        self.cuda_direct = cuda_direct
        if precision not in _native.PREC:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.energy = None
        """Total system energy, valid after a call to **initialise**."""
        self._ncps = 2 ** (self.R - 1)
        self._lib = _native.lib()
        self._ctx = ctypes.c_void_p()
        if not (self._lib.kmc_min_order() <= self.L <= self._lib.kmc_max_order()):
            raise ValueError("l must be in [{}, {}]".format(self._lib.kmc_min_order(), self._lib.kmc_max_order()))
        _native.check(self._lib.kmc_create(ctypes.byref(self._ctx), self.N, self.extent, self.R, self.L,
                                           _native.BC[self._bc.value], self.max_move, self.energy_unit,
                                           int(occ_headroom)), "kmc_create")
        if init_method not in INIT_METHODS:
            raise ValueError("init_method must be one of {}".format(sorted(INIT_METHODS)))
        _native.check(self._lib.kmc_set_init_method(self._ctx, INIT_METHODS[init_method]), "kmc_set_init_method")
        self.profile = {}

    # ------------------------------------------------------------------ helpers
    def _torch(self):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("coulomb_kmc_amd needs a ROCm GPU (torch.cuda is not available)")
        return torch

    def _stream(self):
        torch = self._torch()
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _dev(self, arr, dtype):
        torch = self._torch()
        return torch.as_tensor(np.ascontiguousarray(arr), dtype=dtype).cuda()

    def _pos(self):
        return _as_array(self.positions, 3, np.float64)

    def _q(self):
        return _as_array(self.charges, 1, np.float64).reshape(-1)

    def _assert_init(self):
        if self.energy is None:
            raise RuntimeError("Run initialise before this call")

    def _profile_inc(self, key, inc):
        self.profile[key] = self.profile.get(key, 0.0) + inc

    def profile_kernels(self, on=True):
        """Turns the per-kernel device profile on (or off) and clears it: HIP event pairs
        around the near field, far field, long range and accept sweep of every propose /
        accept called from the host (none under graph capture)."""
        _native.check(self._lib.kmc_profile_enable(self._ctx, int(bool(on))), "kmc_profile_enable")

    def kernel_profile(self):
        """The device profile as the reference's counters (coulomb_kmc/kmc_local.py:466-492,
        kmc_octal.py:176-183, 352-358), also merged into ``profile``: seconds, counts and
        G rates.  The near field evaluates new and old positions in one launch, so
        c_direct_time is the reference's c_direct_new_time + c_direct_old_time and
        c_direct_div_count its two div counts together.  Flop counts: far field 9 nh per
        point, accept sweep 54 images x (9 nh + 6 (L - 1) + 30) per leaf (every order;
        the blocked sweep skips truncated orders, so its rate is an upper count)."""
        out = np.zeros(9)
        _native.check(self._lib.kmc_profile_read(self._ctx, out.ctypes.data_as(ctypes.c_void_p)),
                      "kmc_profile_read")
        nh = self.L * (self.L + 1) // 2
        d = {
            "c_direct_time": out[0] * 1e-3,
            "c_direct_exec_count": int(out[1]),
            "c_direct_div_count": int(out[2]),
            "c_indirect_time": out[3] * 1e-3,
            "c_indirect_exec_count": int(out[4]),
            "c_long_range_time": out[5] * 1e-3,
            "c_long_range_exec_count": int(out[6]),
            "c_accept_time": out[7] * 1e-3,
            "c_accept_exec_count": int(out[8]),
            "c_accept_flop_count_local_create": 9 * nh + 6 * (self.L - 1) + 30,
            "c_accept_exec_count_local_create": int(out[8]) * self.ncells * 54,
        }
        d["c_direct_div_sqrt_grate"] = d["c_direct_div_count"] / d["c_direct_time"] / 1e9 if out[0] > 0 else 0.0
        d["c_indirect_gflop_rate"] = (d["c_indirect_exec_count"] * 9 * nh / d["c_indirect_time"] / 1e9
                                      if out[3] > 0 else 0.0)
        d["c_accept_gflop_rate"] = (d["c_accept_flop_count_local_create"] * d["c_accept_exec_count_local_create"]
                                    / d["c_accept_time"] / 1e9 if out[7] > 0 else 0.0)
        self.profile.update(d)
        return d

    @property
    def handle(self):
        """The C-ABI context (kmc_ctx*)."""
        return self._ctx

    def set_tiles(self, near_lanes=27, sweep_block=0, sweep_pairs=2, sweep_split=1, near_order=0):
        _native.check(self._lib.kmc_set_tiles(self._ctx, int(near_lanes), int(sweep_block), int(sweep_pairs),
                                              int(sweep_split)),
                      "kmc_set_tiles")
        _native.check(self._lib.kmc_set_near_order(self._ctx, int(near_order)), "kmc_set_near_order")

    # ------------------------------------------------------------------ API
    def initialise(self):
        """Initial solve (coulomb_kmc/kmc_fmm.py:947-993); sets ``energy``."""
        t0 = time.time()
        if not self._N_given:
            # the charge set may have changed size since construction (as ppmd dats do)
            self.N = int(self._pos().shape[0])
        n_ctx = ctypes.c_int64(0)
        _native.check(self._lib.kmc_particle_count(self._ctx, ctypes.byref(n_ctx)), "kmc_particle_count")
        if n_ctx.value != self.N:
            _native.check(self._lib.kmc_set_particle_count(self._ctx, self.N), "kmc_set_particle_count")
        pos = np.ascontiguousarray(self._pos()[:self.N], dtype=np.float64)
        q = np.ascontiguousarray(self._q()[:self.N], dtype=np.float64)
        _native.check(self._lib.kmc_set_particles(self._ctx, pos.ctypes.data_as(ctypes.c_void_p),
                                                  q.ctypes.data_as(ctypes.c_void_p)), "kmc_set_particles")
        if self._bc in (BCType.PBC, BCType.FF_ONLY):
            d, p, i = _lr_operator_csr(self.L, self.extent)
            _native.check(self._lib.kmc_set_lr_operator(
                self._ctx, len(d), d.ctypes.data_as(ctypes.c_void_p), p.ctypes.data_as(ctypes.c_void_p),
                i.ctypes.data_as(ctypes.c_void_p)), "kmc_set_lr_operator")
        e = ctypes.c_double(0.0)
        self._torch()
        _native.check(self._lib.kmc_initialise(self._ctx, self._stream(), ctypes.byref(e)), "kmc_initialise")
        self.energy = e.value
        # the per-charge potential of the initial state, as ppmd's FMM writes it into
        # group._kmc_fmm_potential (coulomb_kmc/kmc_fmm.py:942-958); [N, 1], no energy_unit
        self.potential = self.charge_potentials(np.arange(self.N, dtype=np.int64)).reshape(-1, 1)
        self._profile_inc("initialise", time.time() - t0)

    def charge_potentials(self, ids):
        """Potential at each charge ids[n] from all the other charges (near field, leaf
        expansion and, for pbc / ff-only, the long-range term; energy_unit not applied)."""
        ids = np.ascontiguousarray(np.asarray(ids, dtype=np.int64).ravel())
        if ids.size == 0:
            return np.zeros(0)
        torch = self._torch()
        d_ids = torch.as_tensor(ids).cuda()
        d_out = torch.empty(ids.size, dtype=torch.float64, device="cuda")
        _native.check(self._lib.kmc_charge_potentials(self._ctx, self._stream(), ids.size, d_ids.data_ptr(),
                                                      d_out.data_ptr()), "kmc_charge_potentials")
        return d_out.cpu().numpy()

    def _check_move_valid(self, old_pos, new_pos):
        # coulomb_kmc/kmc_mpi_decomp.py:312-345
        e = self.domain.extent
        for dimx in range(3):
            if np.any(new_pos[:, dimx] < -0.5 * e[dimx]) or np.any(new_pos[:, dimx] > 0.5 * e[dimx]):
                raise RuntimeError(
                    "Proposed position is outside simulation domain. Extent {}, dimension {}, value {}.".format(
                        e, dimx, new_pos[:, dimx]))
        rvec = new_pos - old_pos[None, :]
        if self._bc in (BCType.PBC, BCType.NEAREST, BCType.FF_ONLY):
            rvec = np.where(rvec < -0.5 * e, rvec + e, rvec)
            rvec = np.where(rvec > 0.5 * e, rvec - e, rvec)
        if np.any(np.abs(rvec) > self.max_move):
            raise RuntimeError("Proposed move violates max_move. Old position {}, new position {}.".format(
                old_pos, new_pos))

    def propose(self, moves):
        """Total energies of proposed moves ((pid, new_positions), ...); kmc_fmm.py:774-787."""
        self._assert_init()
        t0 = time.time()
        torch = self._torch()
        pos = self._pos()
        pids, es, allpos = [], [0], []
        for pid, mv in moves:
            mv = np.atleast_2d(np.asarray(mv, dtype=np.float64))
            self._check_move_valid(pos[pid], mv)
            pids.append(int(pid))
            es.append(es[-1] + mv.shape[0])
            allpos.append(mv)
        if not pids:
            return tuple()
        T = es[-1]
        d_pid = self._dev(np.array(pids, np.int64), torch.int64)
        d_es = self._dev(np.array(es, np.int64), torch.int64)
        d_pos = self._dev(np.concatenate(allpos), torch.float64)
        d_dE = torch.empty(T, dtype=torch.float64, device="cuda")
        _native.check(self._lib.kmc_propose(self._ctx, self._stream(), len(pids), T, d_pid.data_ptr(), d_es.data_ptr(),
                                            d_pos.data_ptr(), None, d_dE.data_ptr(), _native.PREC[self.precision]),
                      "kmc_propose")
        dE = d_dE.cpu().numpy()
        out = []
        for i, (pid, mv) in enumerate(moves):
            mv = np.atleast_2d(np.asarray(mv, dtype=np.float64))
            vals = self.energy + dE[es[i]:es[i + 1]]
            same = np.linalg.norm(mv - pos[pid][None, :], axis=1) < 1e-14
            vals[same] = self.energy
            out.append(vals)
        self._profile_inc("propose", time.time() - t0)
        self._profile_inc("propose_count", T)
        return tuple(out)

    def propose_with_dats(self, site_max_counts, current_sites, prop_positions, prop_masks, prop_energy_diffs,
                          diff=True, prop_charges=None):
        """Energy differences of masked proposals, written into prop_energy_diffs (kmc_fmm.py:665-762)."""
        self._assert_init()
        t0 = time.time()
        if not diff:
            raise NotImplementedError()
        torch = self._torch()
        masks = _as_array(prop_masks, 1, np.int64)
        diffs = _as_array(prop_energy_diffs, 1, np.float64)
        ppos = _as_array(prop_positions, 3, np.float64)
        assert diffs.shape[1] == masks.shape[1]
        assert diffs.shape[1] * 3 == ppos.shape[1]
        assert diffs.dtype == np.float64
        assert masks.dtype == np.int64
        M = masks.shape[1]
        n = self.N
        smc = np.ascontiguousarray(np.asarray(site_max_counts, dtype=np.int64).ravel())
        sites = np.ascontiguousarray(_as_array(current_sites, 1, np.int64)[:n].ravel())
        if masks.shape[0] < n or ppos.shape[0] < n or diffs.shape[0] < n or sites.shape[0] < n:
            raise RuntimeError("propose_with_dats: the dats hold fewer than N rows")
        if n and (sites.min() < 0 or sites.max() >= smc.size):
            raise RuntimeError("propose_with_dats: a current site is not an index of site_max_counts")
        d_smc = self._dev(smc, torch.int64)
        d_sites = self._dev(sites, torch.int64)
        d_ppos = self._dev(ppos[:n], torch.float64)
        d_masks = self._dev(masks[:n], torch.int64)
        d_out = self._dev(diffs[:n], torch.float64)
        d_pq = None
        if prop_charges is not None:
            d_pq = self._dev(_as_array(prop_charges, 1, np.float64)[:n], torch.float64)
        tm = ctypes.c_int64(0)
        _native.check(self._lib.kmc_propose_with_dats(
            self._ctx, self._stream(), M, d_smc.data_ptr(), smc.size, d_sites.data_ptr(), d_ppos.data_ptr(),
            d_masks.data_ptr(),
            None if d_pq is None else d_pq.data_ptr(), d_out.data_ptr(), _native.PREC[self.precision],
            ctypes.byref(tm)), "kmc_propose_with_dats")
        diffs[:n] = d_out.cpu().numpy()
        self._profile_inc("propose_with_dats", time.time() - t0)
        self._profile_inc("propose_count", tm.value)

    def get_old_energy_with_dats(self, masks, energy):
        raise NotImplementedError()

    def accept(self, move=None):
        """Accept (pid, new_position); moves the charge in ``positions`` (kmc_fmm.py:789-912)."""
        self._assert_init()
        if move is None:
            raise RuntimeError("No moves to accept were passed.")
        t0 = time.time()
        pid = int(move[0])
        newp = np.ascontiguousarray(np.asarray(move[1], dtype=np.float64).reshape(3))
        pos = self._pos()
        self._check_move_valid(pos[pid], newp[None, :])
        e = ctypes.c_double(0.0)
        self._torch()
        _native.check(self._lib.kmc_accept(self._ctx, self._stream(), pid, newp.ctypes.data_as(ctypes.c_void_p),
                                           ctypes.byref(e)), "kmc_accept")
        self.energy = e.value
        pos[pid, :] = newp
        self._profile_inc("accept", time.time() - t0)

    def eval_field(self, points):
        """Potential of all charges at (N, 3) points (coulomb_kmc/kmc_fmm.py:544-567):
        near field, leaf local expansion and (pbc, ff-only) long-range term.
        energy_unit does not affect the output.  pbc potentials are those of the cubic
        lattice sum with conducting boundaries: the zero-mean Ewald potential minus the
        constant (2 pi / 3V) sum_j q_j |x_j|^2."""
        self._assert_init()
        torch = self._torch()
        pts = np.ascontiguousarray(np.atleast_2d(points), dtype=np.float64)
        if pts.shape[1] != 3:
            raise ValueError("points must have shape (N, 3)")
        n = pts.shape[0]
        d_pts = torch.as_tensor(pts).cuda()
        d_out = torch.zeros(n, dtype=torch.float64, device="cuda")
        _native.check(self._lib.kmc_eval_field(self._ctx, self._stream(), n, d_pts.data_ptr(), d_out.data_ptr()),
                      "eval_field")
        return d_out.cpu().numpy()

    def free(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self._lib.kmc_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ------------------------------------------------------------------ state access
    @property
    def ncells(self):
        return self._ncps ** 3

    def get_expansions(self):
        """Leaf local expansions in the reference layout [ncells, 2 L^2]."""
        out = np.zeros((self.ncells, 2 * self.L * self.L))
        _native.check(self._lib.kmc_get_expansions(self._ctx, out.ctypes.data_as(ctypes.c_void_p)), "get_expansions")
        return out

    def set_expansions(self, full):
        full = np.ascontiguousarray(full, dtype=np.float64)
        _native.check(self._lib.kmc_set_expansions(self._ctx, full.ctypes.data_as(ctypes.c_void_p)), "set_expansions")

    def get_multipole(self):
        out = np.zeros(2 * self.L * self.L)
        _native.check(self._lib.kmc_get_multipole(self._ctx, out.ctypes.data_as(ctypes.c_void_p)), "get_multipole")
        return out

    def get_store(self):
        """(occ[ncells], ids[ncells, occmax], pos4[ncells, occmax, 4]) of the cell tables."""
        om = ctypes.c_int32(0)
        _native.check(self._lib.kmc_store_occmax(self._ctx, ctypes.byref(om)), "occmax")
        occ = np.zeros(self.ncells, np.int32)
        ids = np.zeros((self.ncells, om.value), np.int32)
        p4 = np.zeros((self.ncells, om.value, 4))
        _native.check(self._lib.kmc_get_store(self._ctx, occ.ctypes.data_as(ctypes.c_void_p),
                                              ids.ctypes.data_as(ctypes.c_void_p), p4.ctypes.data_as(ctypes.c_void_p),
                                              None), "get_store")
        return occ, ids, p4

    def get_device_particles(self):
        pos = np.zeros((self.N, 3))
        cells = np.zeros(self.N, np.int32)
        _native.check(self._lib.kmc_get_particles(self._ctx, pos.ctypes.data_as(ctypes.c_void_p),
                                                  cells.ctypes.data_as(ctypes.c_void_p)), "get_particles")
        return pos, cells

    def device_energy(self):
        e = ctypes.c_double(0.0)
        _native.check(self._lib.kmc_get_energy(self._ctx, ctypes.byref(e)), "get_energy")
        return e.value

    def leaf_of(self, pos):
        return leaf_of(pos, self.extent, self._ncps)

    # ------------------------------------------------------------------ device-resident path
    def propose_device(self, d_pid, d_es, d_new_pos, d_dE, precision=None, d_new_q=None, stream=None):
        """kmc_propose on torch device tensors (int64 pid[P], es[P+1], float64 new_pos[T,3], dE[T])."""
        P = d_pid.numel()
        T = d_new_pos.numel() // 3
        prec = _native.PREC[precision or self.precision]
        s = stream if stream is not None else self._stream()
        _native.check(self._lib.kmc_propose(self._ctx, s, P, T, d_pid.data_ptr(), d_es.data_ptr(),
                                            d_new_pos.data_ptr(), None if d_new_q is None else d_new_q.data_ptr(),
                                            d_dE.data_ptr(), prec), "kmc_propose")

    def step_device(self, d_pid, d_es, d_new_pos, d_dE, d_uniform, kT, d_counter, d_selected, precision=None,
                    stream=None):
        P = d_pid.numel()
        T = d_new_pos.numel() // 3
        prec = _native.PREC[precision or self.precision]
        s = stream if stream is not None else self._stream()
        nsteps = d_uniform.numel()
        assert d_selected.numel() >= nsteps
        _native.check(self._lib.kmc_step(self._ctx, s, P, T, d_pid.data_ptr(), d_es.data_ptr(), d_new_pos.data_ptr(),
                                         d_dE.data_ptr(), nsteps, d_uniform.data_ptr(), float(kT),
                                         d_counter.data_ptr(), d_selected.data_ptr(), prec), "kmc_step")

    def run_steps(self, nrun, P, k, d_pids_all, d_disp_all, d_uniform, kT, d_counter, d_selected, d_pid, d_es,
                  d_new_pos, d_dE, precision=None, pipelined=True, stream=None):
        """nrun x (materialize_moves + step_device) in one call; pipelined (True or 1): the
        accept sweep of a step overlaps the next step's set-up, near field and long range;
        2: only its set-up and long range; 3: as 2, with the next far field on the sweep's
        stream behind it, beside the near field; 4: near field, combine, select and sweep
        on the call's stream, the tail and the next set-up and long range on a side stream,
        the far field on another."""
        s = stream if stream is not None else self._stream()
        nsteps = d_pids_all.numel() // int(P)
        assert d_disp_all.numel() >= nsteps * int(P) * int(k) * 3
        assert d_selected.numel() >= d_uniform.numel()
        assert d_pid.numel() >= int(P) and d_es.numel() >= int(P) + 1
        assert d_new_pos.numel() >= int(P) * int(k) * 3 and d_dE.numel() >= int(P) * int(k)
        prec = _native.PREC[precision or self.precision]
        _native.check(self._lib.kmc_run_steps(
            self._ctx, s, int(nrun), nsteps, int(P), int(k), d_pids_all.data_ptr(), d_disp_all.data_ptr(),
            d_uniform.numel(), d_uniform.data_ptr(), float(kT), d_counter.data_ptr(), d_selected.data_ptr(),
            d_pid.data_ptr(), d_es.data_ptr(), d_new_pos.data_ptr(), d_dE.data_ptr(), prec,
            int(pipelined)), "kmc_run_steps")

    def materialize_moves(self, P, k, d_pids_all, d_disp_all, d_counter, d_pid, d_es, d_new_pos, stream=None):
        s = stream if stream is not None else self._stream()
        nsteps = d_pids_all.numel() // int(P)
        assert d_disp_all.numel() >= nsteps * int(P) * int(k) * 3
        _native.check(self._lib.kmc_materialize_moves(self._ctx, s, nsteps, int(P), int(k), d_pids_all.data_ptr(),
                                                      d_disp_all.data_ptr(), d_counter.data_ptr(), d_pid.data_ptr(),
                                                      d_es.data_ptr(), d_new_pos.data_ptr()), "materialize")

    def reserve_store(self, slots, stream=None):
        """Gives every leaf at least `slots` free slots for the next device-side run (a run
        of n steps needs n), growing the cell tables when needed
        (coulomb_kmc/kmc_local.py:558-578).  Returns True when they grew: graphs captured
        before must be captured again.  Synchronises."""
        g = ctypes.c_int32(0)
        s = stream if stream is not None else self._stream()
        _native.check(self._lib.kmc_store_reserve(self._ctx, s, int(slots), ctypes.byref(g)), "kmc_store_reserve")
        return bool(g.value)

    def check_flags(self):
        f = ctypes.c_int32(0)
        _native.check(self._lib.kmc_check_flags(self._ctx, self._stream(), ctypes.byref(f)), "check_flags")
        return f.value

    KERNELS = {"sweep": 0, "near": 1, "far": 2, "lr": 3}

    def time_kernel(self, which, reps=20, precision=None, stream=None):
        """Average ms per launch of one kernel (HIP events on the stream); 'sweep' changes the state."""
        ms = ctypes.c_double(0.0)
        s = stream if stream is not None else self._stream()
        _native.check(self._lib.kmc_time_kernel(self._ctx, s, self.KERNELS[which], int(reps),
                                                _native.PREC[precision or self.precision], ctypes.byref(ms)),
                      "time_kernel")
        return ms.value

    def sweep_hop(self, q, old_pos, new_pos, stream=None):
        o = np.ascontiguousarray(old_pos, dtype=np.float64)
        n = np.ascontiguousarray(new_pos, dtype=np.float64)
        s = stream if stream is not None else self._stream()
        _native.check(self._lib.kmc_sweep_hop(self._ctx, s, float(q), o.ctypes.data_as(ctypes.c_void_p),
                                              n.ctypes.data_as(ctypes.c_void_p)), "sweep_hop")
