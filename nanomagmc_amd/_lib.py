"""ctypes binding of libnmc.so (declarations in include/nmc.h).

The library is built in-tree by nanomagmc_amd.build; importing this module
fails loudly if it is missing -- there is no fallback implementation.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnmc.so")

NMC_OK = 0
NMC_E_INVALID = -1
NMC_E_HIP = -2
NMC_E_NOMEM = -3
NMC_E_UNSUPPORTED = -4
NMC_TRIAL_SPHERE = 0
NMC_TRIAL_CONE = 1
NMC_NOBS = 11
# sweep kernel families (nmc_sweep_kinds)
NMC_KIND_STREAM, NMC_KIND_QUADS, NMC_KIND_SPANS, NMC_KIND_LDS, NMC_KIND_LDS_COMPACT, NMC_KIND_COMPACT_STREAM = range(6)
KIND_NAMES = ["stream", "quads", "spans", "lds", "lds_compact", "compact_stream"]

OBS_NAMES = ["Mx", "My", "Mz", "subMx", "subMy", "subMz", "sub4Mx", "sub4My", "sub4Mz", "E_r", "E_s"]

# Every symbol include/nmc.h declares, in header order.
EXPORTS = [
    "nmc_last_error", "nmc_create", "nmc_destroy", "nmc_get_stream", "nmc_sync", "nmc_num_sites", "nmc_num_slices",
    "nmc_set_params", "nmc_set_cone", "nmc_randomize", "nmc_sweep", "nmc_sweep_adapt", "nmc_set_path", "nmc_set_lds_block", "nmc_set_stream_block", "nmc_sweep_kinds", "nmc_get_accepts",
    "nmc_reset_accepts", "nmc_snapshot", "nmc_measure", "nmc_measure_noalloc", "nmc_series_reset", "nmc_sample",
    "nmc_get_series", "nmc_run_samples", "nmc_run_samples_avg", "nmc_accumulate_avg", "nmc_accumulate_avgs", "nmc_get_avg", "nmc_zero_avg", "nmc_get_spins",
    "nmc_set_spins", "nmc_delta_e", "nmc_replay_ising", "nmc_net_create", "nmc_net_destroy",
    "nmc_net_num_params", "nmc_net_get_params", "nmc_net_set_params", "nmc_net_collect",
    "nmc_net_num_samples", "nmc_net_train", "nmc_net_loss_grad", "nmc_block_moves", "nmc_block_swaps", "nmc_get_block_accepts", "nmc_build_weibull",
]


class NmcConfig(C.Structure):
    _fields_ = [("dim", C.c_int), ("ising", C.c_int), ("periodic", C.c_int), ("use_j", C.c_int),
                ("use_d", C.c_int), ("trial", C.c_int), ("seed", C.c_uint64)]


class NmcBond(C.Structure):
    _fields_ = [("d", C.c_int * 3), ("J", C.c_double), ("D", C.c_double * 3)]


class NmcInstanceDesc(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("origin", C.c_int * 3), ("edge", C.c_int),
                ("mask", C.POINTER(C.c_uint8))]


class NmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nmc error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libnmc.so is not built (%s); run python -m nanomagmc_amd.build" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    i, d, u64, vp = C.c_int, C.c_double, C.c_uint64, C.c_void_p
    sig = {
        "nmc_last_error": (C.c_char_p, []),
        "nmc_create": (i, [P(NmcConfig), P(NmcBond), i, P(NmcInstanceDesc), i, P(vp)]),
        "nmc_destroy": (None, [vp]),
        "nmc_get_stream": (i, [vp, P(vp)]),
        "nmc_sync": (i, [vp]),
        "nmc_num_sites": (i, [vp, i, P(C.c_int64)]),
        "nmc_num_slices": (i, [vp, P(i)]),
        "nmc_set_params": (i, [vp, P(i), i, P(d), P(d)]),
        "nmc_set_cone": (i, [vp, P(i), i, P(C.c_float)]),
        "nmc_randomize": (i, [vp, P(i), i]),
        "nmc_sweep": (i, [vp, P(i), i, i]),
        "nmc_sweep_adapt": (i, [vp, P(i), i, i, i, P(C.c_float)]),
        "nmc_set_path": (i, [vp, i]),
        "nmc_set_lds_block": (i, [vp, i]),
        "nmc_get_accepts": (i, [vp, P(i), i, P(u64)]),
        "nmc_reset_accepts": (i, [vp]),
        "nmc_snapshot": (i, [vp, P(i), P(i), i]),
        "nmc_measure": (i, [vp, P(i), i, P(d), P(d), P(d)]),
        "nmc_run_samples_avg": (i, [vp, P(i), i, i, i, P(i), i]),
        "nmc_set_stream_block": (i, [vp, i]),
        "nmc_sweep_kinds": (i, [vp, P(i), i, P(i)]),
        "nmc_measure_noalloc": (i, [vp, P(i), i, P(d), P(d), P(d)]),
        "nmc_series_reset": (i, [vp, i, i]),
        "nmc_sample": (i, [vp, P(i), i, i]),
        "nmc_get_series": (i, [vp, P(i), i, P(d), P(d), P(d)]),
        "nmc_run_samples": (i, [vp, P(i), i, i, i]),
        "nmc_accumulate_avg": (i, [vp, i]),
        "nmc_accumulate_avgs": (i, [vp, P(i), i]),
        "nmc_get_avg": (i, [vp, i, P(d)]),
        "nmc_zero_avg": (i, [vp, i]),
        "nmc_get_spins": (i, [vp, i, P(d)]),
        "nmc_set_spins": (i, [vp, i, P(d)]),
        "nmc_delta_e": (i, [vp, i, P(i), P(d), i, P(d)]),
        "nmc_replay_ising": (i, [vp, i, P(i), P(C.c_int8), P(C.c_uint32), i, P(C.c_uint8)]),
        "nmc_build_weibull": (i, [i, i, P(d), i, d, u64, P(C.c_uint8)]),
        "nmc_net_create": (i, [u64, P(vp)]),
        "nmc_net_destroy": (None, [vp]),
        "nmc_net_num_params": (i, [P(i)]),
        "nmc_net_get_params": (i, [vp, P(C.c_float)]),
        "nmc_net_set_params": (i, [vp, P(C.c_float)]),
        "nmc_net_collect": (i, [vp, vp, i, i, u64]),
        "nmc_net_num_samples": (i, [vp, P(C.c_int64)]),
        "nmc_net_train": (i, [vp, i, i, C.c_float, u64, P(C.c_float)]),
        "nmc_net_loss_grad": (i, [vp, P(C.c_float), P(C.c_float), i, P(C.c_float), P(C.c_float)]),
        "nmc_block_moves": (i, [vp, vp, P(i), i, i, i, P(C.c_int64), P(C.c_float), P(C.c_float)]),
        "nmc_block_swaps": (i, [vp, vp, P(i), i, i, i, i, C.c_float, P(C.c_int64), P(C.c_float), P(C.c_float)]),
        "nmc_get_block_accepts": (i, [vp, P(i), i, P(u64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != NMC_OK:
        raise NmcError(rc, lib().nmc_last_error().decode())
    return rc
