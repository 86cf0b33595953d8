"""FockEngine: the Python face of the MI355X two-electron Fock build.

Mirrors the calls QUICK's Fortran makes around ``gpu_get_cshell_eri`` (the
basis/cutoff uploads once per geometry, then F += G(P) per SCF iteration),
as an object that owns one device-resident engine (``include/quickfock.h``).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .basis import BasisSet


class FockEngine:
    """G(P) = J(P) - hyb/2 K(P) over a Cartesian s/p/d basis on one GPU.

    Screening follows QUICK's CPU get2e: a shell quartet runs when
    Y_ij*Y_kl > cutoff and Y_ij*Y_kl*DNmax > cutoff, a primitive quartet when
    DNmax*cutprim_ij*cutprim_kl > prim_limit (default 0.1*cutoff, as QUICK's
    ``CUTOFF=`` keyword).
    """

    def __init__(self, basis: BasisSet, device: int = 0, cutoff: float = 1e-12,
                 prim_limit: float | None = None, hyb: float = 1.0, precision: str = "fp64",
                 split: float = 1e-4, schwarz=None):
        self.lib = L.lib()
        self.basis = basis
        self.n = basis.nbasis
        h = ctypes.c_void_p()
        L.check(self.lib.qf_engine_create(device, ctypes.byref(h)), "qf_engine_create")
        self.h = h
        mol = basis.mol
        xyz = np.ascontiguousarray(mol.xyz, dtype=np.float64)
        L.check(self.lib.qf_set_molecule(self.h, mol.natom, L.dptr(xyz)), "qf_set_molecule")
        a = basis.shell_arrays()
        self._arrays = a
        L.check(self.lib.qf_set_basis(self.h, len(basis.shells), L.iptr(a["atom"]), L.iptr(a["lmin"]),
                                      L.iptr(a["lmax"]), L.iptr(a["nprim"]), L.iptr(a["pstart"]),
                                      L.iptr(a["fstart"]), basis.nprim, L.dptr(a["exps"]),
                                      L.dptr(np.ascontiguousarray(a["coef"])), basis.nbasis), "qf_set_basis")
        self.set_cutoffs(cutoff, prim_limit, hyb)
        self.set_precision(precision, split)
        if schwarz is not None:
            y, c = (np.ascontiguousarray(x, dtype=np.float64) for x in schwarz)
            L.check(self.lib.qf_set_schwarz(self.h, L.dptr(y), L.dptr(c)), "qf_set_schwarz")
        else:
            null = ctypes.POINTER(ctypes.c_double)()
            L.check(self.lib.qf_set_schwarz(self.h, null, null), "qf_set_schwarz (GPU)")

    def close(self):
        if getattr(self, "h", None):
            self.lib.qf_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def set_cutoffs(self, cutoff, prim_limit=None, hyb=1.0):
        self.cutoff = float(cutoff)
        self.prim_limit = float(0.1 * cutoff if prim_limit is None else prim_limit)
        self.hyb = float(hyb)
        L.check(self.lib.qf_set_cutoffs(self.h, self.cutoff, self.prim_limit, self.hyb), "qf_set_cutoffs")

    def set_precision(self, precision="fp64", split=1e-4):
        mode = {"fp64": L.QF_PREC_FP64, "mixed": L.QF_PREC_MIXED}[precision]
        self.precision = precision
        L.check(self.lib.qf_set_precision(self.h, mode, float(split)), "qf_set_precision")

    def schwarz(self):
        ns = len(self.basis.shells)
        y = np.zeros((ns, ns))
        c = np.zeros((self.basis.nprim, self.basis.nprim))
        L.check(self.lib.qf_get_schwarz(self.h, L.dptr(y), L.dptr(c)), "qf_get_schwarz")
        return y, c

    def prepare(self):
        L.check(self.lib.qf_prepare(self.h), "qf_prepare")

    def stats(self):
        s = L.QfStats()
        L.check(self.lib.qf_get_stats(self.h, ctypes.byref(s)), "qf_get_stats")
        return {k: getattr(s, k) for k, _ in L.QfStats._fields_}

    def one_electron(self):
        n = self.n
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        Z = np.ascontiguousarray(self.basis.mol.Z)
        L.check(self.lib.qf_one_electron(self.h, L.dptr(Z), L.dptr(S), L.dptr(T), L.dptr(V)), "qf_one_electron")
        return S, T, V

    # ------------------------------------------------------------------
    def fock2e(self, P, cutmatrix=None):
        """G(P) for a host density (synchronous)."""
        self.prepare()
        P = np.ascontiguousarray(P, dtype=np.float64)
        G = np.zeros((self.n, self.n))
        cm = None if cutmatrix is None else np.ascontiguousarray(cutmatrix, dtype=np.float64)
        cmp = L.dptr(cm) if cm is not None else ctypes.POINTER(ctypes.c_double)()
        L.check(self.lib.qf_fock_build_host(self.h, L.dptr(P), L.dptr(G), cmp), "qf_fock_build_host")
        return G

    def fock2e_device(self, P_ptr, F_ptr=None, stream=None, cutmat_ptr=None):
        """F += G(P) with device pointers (ints); asynchronous on ``stream``."""
        self.prepare()
        L.check(self.lib.qf_fock_build_device(self.h, ctypes.c_void_p(P_ptr), ctypes.c_void_p(F_ptr or 0),
                                              ctypes.c_void_p(cutmat_ptr or 0), ctypes.c_void_p(stream or 0)),
                "qf_fock_build_device")

    def device_g_ptr(self):
        return self.lib.qf_device_g(self.h)

    def flops(self):
        """Algorithmic FLOPs of a build with the last density (per-class table)."""
        return self.lib.qf_flops_estimate(self.h)

    CLASS_NAMES = ["ss", "ps", "pp", "ds", "dp", "dd"]

    def launches(self):
        """[(name, fp32, ntasks)] of the class launches a build issues."""
        self.prepare()
        out = []
        for k in range(self.lib.qf_launch_count(self.h)):
            cb, ck, f32, nt = (ctypes.c_int() for _ in range(4))
            L.check(self.lib.qf_launch_info(self.h, k, ctypes.byref(cb), ctypes.byref(ck), ctypes.byref(f32),
                                            ctypes.byref(nt)), "qf_launch_info")
            out.append(("(%s|%s)" % (self.CLASS_NAMES[cb.value], self.CLASS_NAMES[ck.value]), bool(f32.value),
                        nt.value))
        return out

    def launch_flops(self):
        n = self.lib.qf_launch_count(self.h)
        f = np.zeros(max(n, 1))
        L.check(self.lib.qf_launch_flops(self.h, L.dptr(f)), "qf_launch_flops")
        return f[:n]

    def launch_counts(self):
        n = self.lib.qf_launch_count(self.h)
        q = (ctypes.c_longlong * max(n, 1))()
        p = (ctypes.c_longlong * max(n, 1))()
        L.check(self.lib.qf_launch_counts(self.h, q, p), "qf_launch_counts")
        return np.array(q[:n]), np.array(p[:n])

    def set_tuning(self, replicas=8, shuffle=True):
        L.check(self.lib.qf_set_tuning(self.h, int(replicas), 1 if shuffle else 0), "qf_set_tuning")

    def set_timing(self, on=True):
        L.check(self.lib.qf_set_timing(self.h, 1 if on else 0), "qf_set_timing")

    def read_timing(self):
        n = self.lib.qf_launch_count(self.h)
        ms = np.zeros(max(n, 1))
        calls = np.zeros(max(n, 1), dtype=np.int32)
        L.check(self.lib.qf_read_timing(self.h, L.dptr(ms), L.iptr(calls)), "qf_read_timing")
        return ms[:n], calls[:n]
