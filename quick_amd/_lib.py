"""ctypes binding of libquickfock.so (the C ABI declared in include/quickfock.h).

The library is built in-tree by ``python -m quick_amd.build``; importing this
module raises if it is missing - there is no CPU fallback for the Fock build.
"""
from __future__ import annotations

import ctypes
import os

LIBPATH = os.environ.get("QUICKFOCK_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                           "libquickfock.so")

_lib = None

QF_OK = 0
QF_PREC_FP64 = 0
QF_PREC_MIXED = 1


class QfStats(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_longlong), ("n_tasks_fp64", ctypes.c_longlong),
                ("n_tasks_fp32", ctypes.c_longlong), ("n_quartets_bound", ctypes.c_longlong),
                ("device_bytes", ctypes.c_longlong)]


class QuickFockError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise QuickFockError(f"{LIBPATH} not found: build it with `python -m quick_amd.build`")
    # PyTorch ships its own HIP runtime under the same soname.  Loading it
    # first makes this library bind to that copy, so the process holds one
    # HIP runtime and torch streams / device pointers are valid here.
    # (Loaded the other way round, torch would bring a second runtime that
    # cannot open the device.)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIBPATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    sig = {
        "qf_last_error": (ctypes.c_char_p, []),
        "qf_engine_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "qf_engine_destroy": (None, [vp]),
        "qf_set_molecule": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "qf_set_basis": (ctypes.c_int, [vp, ctypes.c_int, ip, ip, ip, ip, ip, ip, ctypes.c_int, dp, dp,
                                        ctypes.c_int]),
        "qf_set_cutoffs": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "qf_set_precision": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
        "qf_set_split_classes": (ctypes.c_int, [vp, ctypes.c_int]),
        "qf_set_class_split": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "qf_set_schwarz": (ctypes.c_int, [vp, dp, dp]),
        "qf_get_schwarz": (ctypes.c_int, [vp, dp, dp]),
        "qf_prepare": (ctypes.c_int, [vp]),
        "qf_get_stats": (ctypes.c_int, [vp, ctypes.POINTER(QfStats)]),
        "qf_fock_build_device": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "qf_fock_build_device_noalloc": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "qf_set_grad_cutoff": (ctypes.c_int, [vp, ctypes.c_double]),
        "qf_set_grad_kernel": (ctypes.c_int, [vp, ctypes.c_int]),
        "qf_set_grad_class": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "qf_eri_grad_device": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "qf_eri_grad_host": (ctypes.c_int, [vp, dp, dp, dp]),
        "qf_grad_launch_flops": (ctypes.c_int, [vp, dp, ctypes.POINTER(ctypes.c_longlong),
                                                ctypes.POINTER(ctypes.c_longlong)]),
        "qf_fock_build_host": (ctypes.c_int, [vp, dp, dp, dp]),
        "qf_fock_build_os_device": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp]),
        "qf_fock_build_os_host": (ctypes.c_int, [vp, dp, dp, dp, dp, dp]),
        "qf_device_g": (vp, [vp]),
        "qf_fock_build_batch": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp),
                                               ctypes.POINTER(vp), vp]),
        "qf_one_electron": (ctypes.c_int, [vp, dp, dp, dp, dp]),
        "qf_flops_estimate": (ctypes.c_double, [vp]),
        "qf_launch_count": (ctypes.c_int, [vp]),
        "qf_launch_info": (ctypes.c_int, [vp, ctypes.c_int, ip, ip, ip, ip]),
        "qf_launch_flops": (ctypes.c_int, [vp, dp]),
        "qf_launch_counts": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
        "qf_set_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "qf_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]),
        "qf_batch_create_ext": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp),
                                                ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "qf_batch_destroy": (None, [vp]),
        "qf_batch_size": (ctypes.c_int, [vp]),
        "qf_batch_density": (vp, [vp, ctypes.c_int]),
        "qf_batch_fock": (vp, [vp, ctypes.c_int]),
        "qf_batch_g": (vp, [vp, ctypes.c_int]),
        "qf_batch_set_graph": (ctypes.c_int, [vp, ctypes.c_int]),
        "qf_batch_set_density_host": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "qf_batch_get_host": (ctypes.c_int, [vp, ctypes.c_int, dp, dp]),
        "qf_batch_clear_fock": (ctypes.c_int, [vp, vp]),
        "qf_batch_build": (ctypes.c_int, [vp, vp]),
        "qf_batch_launch_count": (ctypes.c_int, [vp]),
        "qf_batch_launch_info": (ctypes.c_int, [vp, ctypes.c_int, ip, ip, ip, ip]),
        "qf_batch_set_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "qf_batch_read_timing": (ctypes.c_int, [vp, dp, ip]),
        "qf_set_tuning": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "qf_set_tile": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "qf_set_class_tuning": (ctypes.c_int, [vp] + [ctypes.c_int] * 6),
        "qf_set_persistent": (ctypes.c_int, [vp, ctypes.c_int]),
        "qf_set_prim_split": (ctypes.c_int, [vp, ctypes.c_int]),
        "qf_set_class_piece": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "qf_set_class_os": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "qf_read_timing": (ctypes.c_int, [vp, dp, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc, what):
    if rc != QF_OK:
        msg = lib().qf_last_error().decode(errors="replace")
        raise QuickFockError(f"{what} failed (code {rc}): {msg}")


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
