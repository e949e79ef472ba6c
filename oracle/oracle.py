"""
Oracle (test infrastructure only): ctypes wrapper around oracle/libkmc_oracle.so,
the CPU FP64 restatement of the reference's propose/accept path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libkmc_oracle.so")

BC = {"free_space": 0, "27": 1, "pbc": 2, "ff-only": 3}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        L.orc_create.restype = P
        L.orc_create.argtypes = [I64, D, I32, I32, I32, D, D]
        L.orc_destroy.argtypes = [P]
        L.orc_set_threads.argtypes = [I32]
        L.orc_get_threads.restype = I32
        L.orc_set_particles.argtypes = [P, P, P]
        L.orc_set_lr.argtypes = [P, I64, P, P, P, P]
        L.orc_initialise.argtypes = [P, P]
        L.orc_init_lr.argtypes = [P]
        L.orc_set_expansions.argtypes = [P, P]
        L.orc_get_expansions.argtypes = [P, P]
        L.orc_get_store_expansions.argtypes = [P, P]
        L.orc_get_ME.argtypes = [P, P, P]
        L.orc_energy.restype = D
        L.orc_energy.argtypes = [P]
        L.orc_set_energy.argtypes = [P, D]
        L.orc_state_energy.restype = D
        L.orc_state_energy.argtypes = [P]
        L.orc_lr_energy.restype = D
        L.orc_lr_energy.argtypes = [P]
        L.orc_lsd.restype = I32
        L.orc_lsd.argtypes = [P]
        L.orc_occmax.restype = I32
        L.orc_occmax.argtypes = [P]
        L.orc_get_store.argtypes = [P, P, P, P]
        L.orc_get_positions.argtypes = [P, P]
        L.orc_get_cells.argtypes = [P, P]
        L.orc_propose.restype = I32
        L.orc_propose.argtypes = [P, I64, P, P, P, P, P, P]
        L.orc_propose_check.argtypes = [P, I64, P, P, P, P, P, P, I32]
        L.orc_propose_with_dats.restype = I32
        L.orc_propose_with_dats.argtypes = [P, I64, P, P, P, P, P, P]
        L.orc_accept.restype = I32
        L.orc_accept.argtypes = [P, I64, P, P]
        L.orc_accept_sweep_only.argtypes = [P, D, P, P]
        L.orc_local_eval.argtypes = [I32, P, P, P]
        L.orc_p2l.argtypes = [I32, D, P, P]
        L.orc_p2m.argtypes = [I32, D, P, P]
        L.orc_dot_vec.argtypes = [I32, D, P, P]
        L.orc_charge_energy.argtypes = [P, I64, P, P]
        L.orc_point_energy.argtypes = [P, I64, P, P, P]
        L.orc_group_energy.restype = D
        L.orc_group_energy.argtypes = [P, I64, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


class Oracle:
    """One system held by the CPU restatement (reference data structures)."""

    def __init__(self, positions, charges, extent, r, L, bc="pbc", max_move=None,
                 energy_unit=1.0, lr=None):
        self.pos0 = _f64(positions).reshape(-1, 3)
        self.q0 = _f64(charges).ravel()
        self.N = self.pos0.shape[0]
        self.E = float(extent)
        self.L = int(L)
        self.r = int(r)
        self.ncps = 2 ** (self.r - 1)
        self.bc = bc
        mm = float(max_move) if max_move is not None else self.E
        self.h = lib().orc_create(self.N, self.E, self.r, self.L, BC[bc], mm, float(energy_unit))
        lib().orc_set_particles(self.h, _p(self.pos0), _p(self.q0))
        if lr is not None:
            data, indptr, indices, dip = lr
            self._lr = (_f64(data), _i64(indptr), _i64(indices), _f64(dip))
            lib().orc_set_lr(self.h, len(self._lr[0]), _p(self._lr[0]), _p(self._lr[1]), _p(self._lr[2]),
                             _p(self._lr[3]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def initialise(self):
        e = np.zeros(1)
        lib().orc_initialise(self.h, _p(e))
        return float(e[0])

    def init_from_expansions(self, full, energy):
        """Take leaf expansions (reference layout) and the energy from elsewhere; build M/E."""
        self.set_expansions(full)
        lib().orc_init_lr(self.h)
        self.energy = energy

    @property
    def energy(self):
        return float(lib().orc_energy(self.h))

    @energy.setter
    def energy(self, v):
        lib().orc_set_energy(self.h, float(v))

    def state_energy(self):
        return float(lib().orc_state_energy(self.h))

    def lr_energy(self):
        return float(lib().orc_lr_energy(self.h))

    def set_expansions(self, full):
        full = _f64(full)
        lib().orc_set_expansions(self.h, _p(full))

    def get_expansions(self):
        out = np.zeros((self.ncps ** 3, 2 * self.L * self.L))
        lib().orc_get_expansions(self.h, _p(out))
        return out

    def get_store_expansions(self):
        lsd = lib().orc_lsd(self.h)
        out = np.zeros((lsd ** 3, 2 * self.L * self.L))
        lib().orc_get_store_expansions(self.h, _p(out))
        return out

    def get_ME(self):
        M = np.zeros(2 * self.L * self.L)
        E = np.zeros(2 * self.L * self.L)
        lib().orc_get_ME(self.h, _p(M), _p(E))
        return M, E

    def get_store(self):
        om = lib().orc_occmax(self.h)
        nc = self.ncps ** 3
        occ = np.zeros(nc, np.int64)
        ids = np.zeros((nc, om), np.int64)
        data = np.zeros((nc, om, 5))
        lib().orc_get_store(self.h, _p(occ), _p(ids), _p(data))
        return occ, ids, data

    def positions(self):
        out = np.zeros((self.N, 3))
        lib().orc_get_positions(self.h, _p(out))
        return out

    def cells(self):
        out = np.zeros(self.N, np.int64)
        lib().orc_get_cells(self.h, _p(out))
        return out

    def propose_arrays(self, pids, es, new_pos, new_q=None, comps=False):
        pids = _i64(pids)
        es = _i64(es)
        new_pos = _f64(new_pos).reshape(-1, 3)
        T = int(es[-1])
        dE = np.zeros(T)
        cm = np.zeros((T, 5)) if comps else None
        nq = None if new_q is None else _f64(new_q)
        err = lib().orc_propose(self.h, len(pids), _p(pids), _p(es), _p(new_pos), _p(nq), _p(dE), _p(cm))
        if err:
            raise RuntimeError("Bad proposed move was detected.")
        return (dE, cm) if comps else dE

    def propose_check(self, pids, es, new_pos, new_q=None, extended=True, comps=False):
        """dE of proposals of the stored state, evaluated in long double (extended) or
        double, with the long-range term in the expanded difference form (no subtraction
        of two whole-system energies): the reference the GPU paths are measured against.
        No validation of the moves."""
        pids = _i64(pids)
        es = _i64(es)
        new_pos = _f64(new_pos).reshape(-1, 3)
        T = int(es[-1])
        dE = np.zeros(T)
        cm = np.zeros((T, 5)) if comps else None
        nq = None if new_q is None else _f64(new_q)
        lib().orc_propose_check(self.h, len(pids), _p(pids), _p(es), _p(new_pos), _p(nq), _p(dE), _p(cm),
                                1 if extended else 0)
        return (dE, cm) if comps else dE

    def propose(self, moves):
        """Tuple API of the reference: returns energy + dE per proposal."""
        pids, es, pos = [], [0], []
        for pid, mv in moves:
            mv = np.atleast_2d(np.asarray(mv, dtype=np.float64))
            pids.append(int(pid))
            es.append(es[-1] + mv.shape[0])
            pos.append(mv)
        dE = self.propose_arrays(pids, es, np.concatenate(pos))
        e = self.energy
        cur = self.positions()
        out = []
        for i, (pid, mv) in enumerate(moves):
            mv = np.atleast_2d(np.asarray(mv, dtype=np.float64))
            vals = np.zeros(mv.shape[0])
            for k in range(mv.shape[0]):
                if np.linalg.norm(cur[pid] - mv[k]) < 1e-14:
                    vals[k] = e
                else:
                    vals[k] = e + dE[es[i] + k]
            out.append(vals)
        return tuple(out)

    def propose_with_dats(self, site_max_counts, current_sites, prop_positions, prop_masks,
                          prop_energy_diffs, prop_charges=None):
        masks = _i64(prop_masks)
        M = masks.shape[1]
        out = np.array(prop_energy_diffs, dtype=np.float64, copy=True)
        pc = None if prop_charges is None else _f64(prop_charges)
        err = lib().orc_propose_with_dats(self.h, M, _p(_i64(site_max_counts)), _p(_i64(current_sites)),
                                          _p(_f64(prop_positions)), _p(masks), _p(pc), _p(out))
        if err:
            raise RuntimeError("Bad proposed move was detected.")
        prop_energy_diffs[...] = out
        return out

    def accept(self, pid, new_pos):
        dE = np.zeros(1)
        err = lib().orc_accept(self.h, int(pid), _p(_f64(new_pos)), _p(dE))
        if err:
            raise RuntimeError("accept failed ({})".format(err))
        return float(dE[0])

    # inject / extract (coulomb_kmc/kmc_inject_extract.py:82-237)
    def charge_energy(self, ids):
        """Per-charge energies (get_energy before the row sum), eu applied."""
        ids = _i64(ids).ravel()
        out = np.zeros(ids.size)
        lib().orc_charge_energy(self.h, ids.size, _p(ids), _p(out))
        return out

    def point_energy(self, pos, q):
        """q_i phi(x_i) at points holding no charge of the system, eu applied."""
        pos = _f64(pos).reshape(-1, 3)
        q = _f64(q).ravel()
        out = np.zeros(q.size)
        lib().orc_point_energy(self.h, q.size, _p(pos), _p(q), _p(out))
        return out

    def group_energy(self, pos, q):
        """compute_energy: the charges alone in the domain, eu applied."""
        pos = _f64(pos).reshape(-1, 3)
        q = _f64(q).ravel()
        return float(lib().orc_group_energy(self.h, q.size, _p(pos), _p(q)))

    def propose_extract(self, ids, self_energy=True):
        """kmc_inject_extract.py:178-213 for rows of ids."""
        ids = np.atleast_2d(_i64(ids))
        cur = self.positions()
        out = np.zeros(ids.shape[0])
        for r, row in enumerate(ids):
            ge = self.group_energy(cur[row], self.q0[row]) if self_energy else 0.0
            out[r] = ge - self.charge_energy(row).sum()
        return out

    def propose_inject(self, pos, q):
        """kmc_inject_extract.py:216-237."""
        return self.group_energy(pos, q) + float(self.point_energy(pos, q).sum())

    def accept_sweep_only(self, q, new_pos, old_pos):
        lib().orc_accept_sweep_only(self.h, float(q), _p(_f64(new_pos)), _p(_f64(old_pos)))


# ---------------------------------------------------------------------------
# the one error measure of proposal energies used by the tests and bench.py
# ---------------------------------------------------------------------------
# Relative error with the reference magnitude floored at DE_FLOOR (energy units,
# = 1e-3 kT at kT = 1): a proposal whose |dE| is below the floor is held to an
# absolute error of bar x DE_FLOOR (1e-13 for the FP64 bar 1e-10, 1e-11 for the
# mixed bar 1e-8).  An absolute error d in dE changes its Metropolis rate by a factor
# exp(-d / kT); below the floor the relative error of dE itself says nothing about
# the rate, while the absolute bound still does.
DE_FLOOR = 1e-3
DE_BAR_FP64 = 1e-10
DE_BAR_MIXED = 1e-8


def de_rel_err(a, ref, floor=DE_FLOOR):
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.abs(a - ref) / np.maximum(np.abs(ref), floor)


def de_error(a, ref, floor=DE_FLOOR):
    """max / median absolute and floored relative error of dE values a against ref."""
    a = np.asarray(a, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    ab = np.abs(a - ref)
    rel = de_rel_err(a, ref, floor)
    return {"proposals": int(a.size), "max_abs": float(ab.max()), "median_abs": float(np.median(ab)),
            "max_rel": float(rel.max()), "median_rel": float(np.median(rel)),
            "rel_floor": floor, "below_floor": int(np.sum(np.abs(ref) < floor))}


# basis primitives (reference layout)
def local_eval(L, lexp, disp):
    out = np.zeros(1)
    lib().orc_local_eval(int(L), _p(_f64(lexp)), _p(_f64(disp)), _p(out))
    return float(out[0])


def p2l(L, q, disp):
    out = np.zeros(2 * L * L)
    lib().orc_p2l(int(L), float(q), _p(_f64(disp)), _p(out))
    return out


def p2m(L, q, disp):
    out = np.zeros(2 * L * L)
    lib().orc_p2m(int(L), float(q), _p(_f64(disp)), _p(out))
    return out


def dot_vec(L, q, disp):
    out = np.zeros(2 * L * L)
    lib().orc_dot_vec(int(L), float(q), _p(_f64(disp)), _p(out))
    return out
