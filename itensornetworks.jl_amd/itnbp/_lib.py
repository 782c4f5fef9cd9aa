"""ctypes binding of libitnbp.so (the C ABI declared in include/itnbp.h).

The library is required: importing the compute entry points without it raises
``ITNBPLibraryMissing``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import atexit
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ITNBP_LIB", os.path.join(HERE, "libitnbp.so"))

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_NONFINITE, ERR_UNSUPPORTED, ERR_STATE = range(7)
F64, F32, BF16 = 0, 1, 2
SCHED_SEQUENTIAL, SCHED_PARALLEL = 0, 1
(PREC_F64, PREC_F32, PREC_BF16_F64, PREC_BF16X3, PREC_BF16X3_F64, PREC_BF16_BF16X3, PREC_BF16_BF16X3_F64,
 PREC_F16X2, PREC_BF16_F16X2) = range(9)

_ERRNAMES = {ERR_ARG: "invalid argument", ERR_HIP: "HIP error", ERR_NOMEM: "out of memory",
             ERR_NONFINITE: "non-finite result", ERR_UNSUPPORTED: "unsupported",
             ERR_STATE: "invalid state"}


class ITNBPLibraryMissing(RuntimeError):
    pass


class ITNBPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class UpdateOpts(C.Structure):
    _fields_ = [("schedule", C.c_int), ("maxiter", C.c_int), ("tol", C.c_double),
                ("res_tol", C.c_double), ("precision", C.c_int), ("switch_tol", C.c_double),
                ("normalize", C.c_int), ("nseq", C.c_int), ("seq", C.POINTER(C.c_int)),
                ("verbose", C.c_int)]


class UpdateResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("iterations_low", C.c_int), ("mean_diff", C.c_double),
                ("max_residual", C.c_double), ("converged", C.c_int), ("device_ms", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/itnbp.h
SIGNATURES = {
    "itnbp_create": (_I, [_I, C.POINTER(_P)]),
    "itnbp_destroy": (_I, [_P]),
    "itnbp_last_error": (C.c_char_p, [_P]),
    "itnbp_synchronize": (_I, [_P]),
    "itnbp_net_create": (_I, [_P, _I, _I, _IP, _IP, _IP, _IP, _IP, _I, C.POINTER(_P)]),
    "itnbp_net_destroy": (_I, [_P]),
    "itnbp_net_copy": (_I, [_P, C.POINTER(_P)]),
    "itnbp_net_psi_numel": (C.c_int64, [_P]),
    "itnbp_get_bond_dims": (_I, [_P, _IP]),
    "itnbp_net_msg_numel": (C.c_int64, [_P]),
    "itnbp_set_sites": (_I, [_P, _P]),
    "itnbp_get_sites": (_I, [_P, _P]),
    "itnbp_set_site": (_I, [_P, _I, _P]),
    "itnbp_set_messages": (_I, [_P, _DP]),
    "itnbp_get_messages": (_I, [_P, _DP]),
    "itnbp_set_message": (_I, [_P, _I, _DP]),
    "itnbp_get_message": (_I, [_P, _I, _DP]),
    "itnbp_init_identity_messages": (_I, [_P]),
    "itnbp_message_is_set": (_I, [_P, _I]),
    "itnbp_forest_cover_sequence": (_I, [_P, _IP]),
    "itnbp_update": (_I, [_P, C.POINTER(UpdateOpts), C.POINTER(UpdateResult)]),
    "itnbp_updated_messages": (_I, [_P, _I, _IP, _I, _DP]),
    "itnbp_rdm1": (_I, [_P, _I, _IP, _DP]),
    "itnbp_einsum": (_I, [_P, _I, _IP, _IP, _IP, _I, _DP, _I, _IP, _DP]),
    "itnbp_expect1": (_I, [_P, _I, _IP, _DP, _DP]),
    "itnbp_rdm2": (_I, [_P, _I, _IP, _DP]),
    "itnbp_expect2": (_I, [_P, _I, _IP, _DP, _DP]),
    "itnbp_region_scalars": (_I, [_P, _DP, _DP]),
    "itnbp_logscalar": (_I, [_P, _DP, _DP, _IP]),
    "itnbp_normalize": (_I, [_P]),
    "itnbp_apply_gates": (_I, [_P, _I, _IP, _DP, _I, _D, _I, _DP, _DP, _I, _IP]),
    "itnbp_apply_site_gates": (_I, [_P, _I, _IP, _DP, _I, _DP]),
    "itnbp_apply_gates_noalloc": (_I, [_P, _I, _IP, _DP, _I, _D, _I, _DP, _DP, _I, _IP]),
    "itnbp_sweeps_vjp": (_I, [_P, _I, _I, _DP, _DP]),
    "itnbp_apply_vjp": (_I, [_P, _I, _IP, _DP, _I, _D, _I, _IP, _DP, _DP, _DP, _DP]),
    "itnbp_energy_grad": (_I, [_P, _I, _I, _DP, _DP, _DP, _DP]),
    "itnbp_update_batched": (_I, [_I, C.POINTER(_P), C.POINTER(UpdateOpts), C.POINTER(UpdateResult)]),
    "itnbp_lnet_create": (_I, [_P, _I, _I, _IP, _IP, _IP, _IP, _IP, _IP, _I, C.POINTER(_P)]),
    "itnbp_lnet_destroy": (_I, [_P]),
    "itnbp_lnet_msg_numel": (C.c_int64, [_P]),
    "itnbp_lnet_ket_numel": (C.c_int64, [_P]),
    "itnbp_lnet_bra_numel": (C.c_int64, [_P]),
    "itnbp_lnet_set_tensors": (_I, [_P, _DP, _DP]),
    "itnbp_lnet_fold_operator": (_I, [_P, _DP, _IP, _DP, _IP]),
    "itnbp_lnet_get_ket": (_I, [_P, _DP]),
    "itnbp_lnet_set_message": (_I, [_P, _I, _DP]),
    "itnbp_lnet_get_message": (_I, [_P, _I, _DP]),
    "itnbp_lnet_message_is_set": (_I, [_P, _I]),
    "itnbp_lnet_init_identity_messages": (_I, [_P]),
    "itnbp_lnet_update": (_I, [_P, C.POINTER(UpdateOpts), C.POINTER(UpdateResult)]),
    "itnbp_lnet_region_scalars": (_I, [_P, _DP, _DP]),
    "itnbp_lnet_logscalar": (_I, [_P, _DP, _DP]),
    "itnbp_lnet_rdm1": (_I, [_P, _I, _IP, _DP]),
    "itnbp_last_path": (C.c_char_p, [_P]),
    "itnbp_profile_enable": (_I, [_P, _I]),
    "itnbp_profile_reset": (_I, [_P]),
    "itnbp_profile_query": (_I, [_P, C.c_char_p, _IP, _DP]),
}

_lib = None


def load(path=None):
    """Load libitnbp.so and declare every entry point."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ITNBPLibraryMissing(
            f"{p} not found: build it with `python itensornetworks.jl_amd/build.py` "
            "(there is no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def iptr(a):
    return ptr(a, C.c_int)


def dptr(a):
    return ptr(a, C.c_double)


def vptr(a):
    return C.c_void_p(a.ctypes.data)


class Context:
    """One HIP device context (a stream and events) of the library."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.itnbp_create(device, C.byref(h))
        if rc != OK:
            raise ITNBPError(rc, f"itnbp_create(device={device}) failed (no usable GPU?)")
        self.h = h
        self.device = device
        # networks created on this context: released before the context itself
        # (at close() or interpreter exit), whatever order the collector picks
        self._nets = weakref.WeakSet()

    def check(self, rc):
        if rc != OK:
            msg = self.lib.itnbp_last_error(self.h)
            raise ITNBPError(rc, msg.decode() if msg else "")

    def profile(self, enable=True):
        self.check(self.lib.itnbp_profile_enable(self.h, int(enable)))
        self.check(self.lib.itnbp_profile_reset(self.h))

    def kernel_time(self, name):
        """(launch count, summed ms) of a library kernel since the last profile reset."""
        n, t = C.c_int(), C.c_double()
        self.check(self.lib.itnbp_profile_query(self.h, name.encode(), C.byref(n), C.byref(t)))
        return n.value, t.value

    def last_path(self):
        return self.lib.itnbp_last_path(self.h).decode()

    def close(self):
        """Release every network of this context, then the context."""
        for net in list(getattr(self, "_nets", ())):
            net._release()
        if getattr(self, "h", None):
            self.lib.itnbp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


@atexit.register
def _close_default_context():
    if _default_ctx is not None:
        _default_ctx.close()


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("ITNBP_DEVICE", "0")))
    return _default_ctx


def as_contig(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)
