"""
Loader and build recipe of the HIP library (libkmc_amd.so, C ABI in include/kmc_amd.h).

The library is built in-tree by ``build()`` (hipcc for gfx950) and loaded with
ctypes.  There is no fallback: if the library is missing, ``lib()`` raises.
"""

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_CSRC = os.path.join(_HERE, "csrc")
_SRC = os.path.join(_CSRC, "kmc_amd.hip")          # order-independent kernels + host code
_ORDER_SRC = os.path.join(_CSRC, "kmc_order.hip")  # compiled once per expansion order
ORDERS = tuple(range(2, 21))
LIB_PATH = os.path.join(_HERE, "libkmc_amd.so")
_OBJ_DIR = os.path.join(_HERE, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wno-unused-value",
    "-Wno-unused-function",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

BC = {"free_space": 0, "27": 1, "pbc": 2, "ff-only": 3}
PREC = {"fp64": 0, "mixed": 1}

KMC_E_ARG = 1
KMC_E_BAD_MOVE = 4
KMC_E_STORE = 5
KMC_E_STATE = 3

# exported symbols (each declared in include/kmc_amd.h)
EXPORTS = (
    "kmc_last_error", "kmc_min_order", "kmc_max_order", "kmc_create", "kmc_destroy",
    "kmc_set_particles", "kmc_set_lr_operator", "kmc_initialise", "kmc_set_init_method", "kmc_get_expansions",
    "kmc_set_expansions", "kmc_get_multipole", "kmc_get_store", "kmc_store_occmax",
    "kmc_get_particles", "kmc_get_energy", "kmc_set_energy", "kmc_ncells", "kmc_propose",
    "kmc_propose_with_dats", "kmc_accept", "kmc_step", "kmc_materialize_moves",
    "kmc_check_flags", "kmc_store_reserve", "kmc_set_tiles", "kmc_set_near_order", "kmc_profile_enable", "kmc_profile_read", "kmc_sweep_hop", "kmc_time_kernel", "kmc_probe_peaks", "kmc_eval_field",
    "kmc_batch_create", "kmc_batch_destroy", "kmc_batch_set_system", "kmc_batch_prepare", "kmc_batch_step", "kmc_batch_time_sweep", "kmc_batch_set_streams", "kmc_run_steps",
    "kmc_particle_count", "kmc_set_particle_count", "kmc_charge_energies", "kmc_charge_potentials", "kmc_group_energy",
    "kmc_propose_extract", "kmc_propose_inject", "kmc_extract", "kmc_inject", "kmc_discover_sites",
    "kmc_lattice_step", "kmc_accept_movedata", "kmc_propose_noalloc",
)


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _compile_jobs():
    jobs = [([_SRC], os.path.join(_OBJ_DIR, "kmc_amd.o"))]
    for L in ORDERS:
        jobs.append((["-DKMC_ORDER={}".format(L), _ORDER_SRC], os.path.join(_OBJ_DIR, "kmc_order_{}.o".format(L))))
    return jobs


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        msg = "\n".join(l for l in res.stderr.splitlines() if "not a recognized feature" not in l)
        raise RuntimeError("hipcc failed: {}\n{}".format(" ".join(cmd[-2:]), msg[-4000:]))


def build(verbose=False, jobs=None):
    """Compile csrc/ into libkmc_amd.so (always recompiles): kmc_amd.hip and one object
    of kmc_order.hip per expansion order, in parallel, then one shared-library link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(_OBJ_DIR, exist_ok=True)
    cc = hipcc()
    todo = _compile_jobs()
    # the order-independent object is the longest single compile: start it first
    nproc = jobs or max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=nproc) as ex:
        futs = [ex.submit(_run, [cc] + HIPCC_FLAGS + ["-c"] + args[:-1] + ["-o", obj, args[-1]])
                for args, obj in todo]
        for f in futs:
            f.result()
    tmp = LIB_PATH + ".tmp"
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [obj for _, obj in todo])
    os.replace(tmp, LIB_PATH)
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


_lib = None


class KMCError(RuntimeError):
    pass


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "coulomb_kmc_amd: HIP library {} is not built; run __graft_entry__.build() "
            "or coulomb_kmc_amd._native.build()".format(LIB_PATH))
    L = ctypes.CDLL(LIB_PATH)
    P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    sig = {
        "kmc_last_error": (ctypes.c_char_p, []),
        "kmc_min_order": (I32, []),
        "kmc_max_order": (I32, []),
        "kmc_create": (ctypes.c_int, [ctypes.POINTER(P), I64, D, I32, I32, I32, D, D, I32]),
        "kmc_destroy": (ctypes.c_int, [P]),
        "kmc_set_particles": (ctypes.c_int, [P, P, P]),
        "kmc_set_lr_operator": (ctypes.c_int, [P, I64, P, P, P]),
        "kmc_initialise": (ctypes.c_int, [P, P, P]),
        "kmc_set_init_method": (ctypes.c_int, [P, I32]),
        "kmc_get_expansions": (ctypes.c_int, [P, P]),
        "kmc_set_expansions": (ctypes.c_int, [P, P]),
        "kmc_get_multipole": (ctypes.c_int, [P, P]),
        "kmc_get_store": (ctypes.c_int, [P, P, P, P, P]),
        "kmc_store_occmax": (ctypes.c_int, [P, P]),
        "kmc_get_particles": (ctypes.c_int, [P, P, P]),
        "kmc_get_energy": (ctypes.c_int, [P, P]),
        "kmc_set_energy": (ctypes.c_int, [P, D]),
        "kmc_ncells": (ctypes.c_int, [P, P]),
        "kmc_propose": (ctypes.c_int, [P, P, I64, I64, P, P, P, P, P, I32]),
        "kmc_propose_noalloc": (ctypes.c_int, [P, P, I64, I64, P, P, P, P, P, I32]),
        "kmc_propose_with_dats": (ctypes.c_int, [P, P, I64, P, I64, P, P, P, P, P, I32, P]),
        "kmc_accept": (ctypes.c_int, [P, P, I64, P, P]),
        "kmc_step": (ctypes.c_int, [P, P, I64, I64, P, P, P, P, I64, P, D, P, P, I32]),
        "kmc_materialize_moves": (ctypes.c_int, [P, P, I64, I64, I64, P, P, P, P, P, P]),
        "kmc_check_flags": (ctypes.c_int, [P, P, P]),
        "kmc_store_reserve": (ctypes.c_int, [P, P, I32, P]),
        "kmc_eval_field": (ctypes.c_int, [P, P, I64, P, P]),
        "kmc_batch_create": (ctypes.c_int, [ctypes.POINTER(P), P, I32]),
        "kmc_batch_destroy": (ctypes.c_int, [P]),
        "kmc_batch_set_system": (ctypes.c_int, [P, I32, I64, I64, I64, P, P, P, P, P, P, P, P, P]),
        "kmc_batch_prepare": (ctypes.c_int, [P, D, I32]),
        "kmc_batch_step": (ctypes.c_int, [P, P]),
        "kmc_batch_time_sweep": (ctypes.c_int, [P, P, I32, P]),
        "kmc_batch_set_streams": (ctypes.c_int, [P, I32]),
        "kmc_run_steps": (ctypes.c_int, [P, P, I64, I64, I64, I64, P, P, I64, P, D, P, P, P, P, P, P, I32, I32]),
        "kmc_set_tiles": (ctypes.c_int, [P, I32, I32, I32, I32]),
        "kmc_set_near_order": (ctypes.c_int, [P, I32]),
        "kmc_profile_enable": (ctypes.c_int, [P, I32]),
        "kmc_profile_read": (ctypes.c_int, [P, P]),
        "kmc_sweep_hop": (ctypes.c_int, [P, P, D, P, P]),
        "kmc_time_kernel": (ctypes.c_int, [P, P, I32, I32, I32, P]),
        "kmc_probe_peaks": (ctypes.c_int, [P, I64, I32, P]),
        "kmc_particle_count": (ctypes.c_int, [P, P]),
        "kmc_set_particle_count": (ctypes.c_int, [P, I64]),
        "kmc_charge_energies": (ctypes.c_int, [P, P, I64, P, P]),
        "kmc_charge_potentials": (ctypes.c_int, [P, P, I64, P, P]),
        "kmc_group_energy": (ctypes.c_int, [P, P, I64, I64, P, P, P]),
        "kmc_propose_extract": (ctypes.c_int, [P, P, I64, I64, P, I32, P]),
        "kmc_propose_inject": (ctypes.c_int, [P, P, I64, P, P, P]),
        "kmc_extract": (ctypes.c_int, [P, P, I64, P, P]),
        "kmc_inject": (ctypes.c_int, [P, P, I64, P, P, P]),
        "kmc_discover_sites": (ctypes.c_int, [P, I64, P, I64, P, I64, P, D, P, P]),
        "kmc_lattice_step": (ctypes.c_int, [P, P, I64, P, D, I32, P, P, P, I64, P, D, I32, P, P, I32]),
        "kmc_accept_movedata": (ctypes.c_int, [P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, what=""):
    """Turn a C-ABI status into the reference's exception types."""
    if rc == 0:
        return
    msg = lib().kmc_last_error().decode(errors="replace")
    if rc == KMC_E_BAD_MOVE:
        raise RuntimeError(msg or "Bad proposed move was detected.")
    if rc == KMC_E_STATE:
        raise RuntimeError(msg or "Run initialise before this call")
    raise KMCError("{} failed ({}): {}".format(what, rc, msg))
