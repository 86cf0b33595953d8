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
    """G(P) = J(P) - hyb/2 K(P) over a Cartesian s/p/d/f basis on one GPU.

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

    def set_split_classes(self, min_lsum=-1):
        """Mixed builds split the classes with La+Lb+Lc+Ld >= min_lsum (-1: by molecule size)."""
        L.check(self.lib.qf_set_split_classes(self.h, int(min_lsum)), "qf_set_split_classes")

    def set_class_split(self, cb, ck, mode=-1):
        """Mixed builds, class (cb|ck): -1 as set_split_classes, 0 all FP64, 1 split."""
        L.check(self.lib.qf_set_class_split(self.h, int(cb), int(ck), int(mode)), "qf_set_class_split")

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

    def fock2e_device_noalloc(self, P_ptr, F_ptr=None, stream=None, cutmat_ptr=None):
        """As fock2e_device, with the build's scratch in the library's fixed
        device array (qf_fock_build_device_noalloc): no device allocation in
        the call once the engine is prepared."""
        self.prepare()
        L.check(self.lib.qf_fock_build_device_noalloc(self.h, ctypes.c_void_p(P_ptr), ctypes.c_void_p(F_ptr or 0),
                                                      ctypes.c_void_p(cutmat_ptr or 0),
                                                      ctypes.c_void_p(stream or 0)),
                "qf_fock_build_device_noalloc")

    def set_grad_cutoff(self, grad_cutoff):
        """QUICK's gradCutoff for the two-electron gradient (default: the
        integral cutoff)."""
        L.check(self.lib.qf_set_grad_cutoff(self.h, float(grad_cutoff)), "qf_set_grad_cutoff")

    def set_grad_kernel(self, lane=True):
        """Per-lane gradient kernels for the light classes (default), or the
        wave-per-quartet kernel for every class."""
        L.check(self.lib.qf_set_grad_kernel(self.h, 1 if lane else 0), "qf_set_grad_kernel")

    def set_grad_class(self, cb, ck, mode=-1):
        """Gradient kernel of class (cb|ck): -1 as set_grad_kernel, 0 the
        wave-per-quartet kernel, 1 the per-lane kernel where the class has one."""
        L.check(self.lib.qf_set_grad_class(self.h, int(cb), int(ck), int(mode)), "qf_set_grad_class")

    def eri_grad(self, P, cutmatrix=None):
        """d/dR of 1/2 sum P G(P) with P fixed, (natom, 3) in Eh/bohr (QUICK's
        gpu_get_cshell_eri_grad_); synchronous."""
        self.prepare()
        P = np.ascontiguousarray(P, dtype=np.float64)
        g = np.zeros((self.basis.mol.natom, 3))
        cm = None if cutmatrix is None else np.ascontiguousarray(cutmatrix, dtype=np.float64)
        cmp = L.dptr(cm) if cm is not None else ctypes.POINTER(ctypes.c_double)()
        L.check(self.lib.qf_eri_grad_host(self.h, L.dptr(P), L.dptr(g), cmp), "qf_eri_grad_host")
        return g

    def eri_grad_device(self, P_ptr, grad_ptr, stream=None, cutmat_ptr=None):
        """grad (3 natom doubles on the device) += the two-electron gradient;
        asynchronous on ``stream``."""
        self.prepare()
        L.check(self.lib.qf_eri_grad_device(self.h, ctypes.c_void_p(P_ptr), ctypes.c_void_p(grad_ptr),
                                            ctypes.c_void_p(cutmat_ptr or 0), ctypes.c_void_p(stream or 0)),
                "qf_eri_grad_device")

    def grad_launch_counts(self):
        """Per Fock launch of the last gradient: (FLOPs, quartets, primitive quartets)."""
        n = self.lib.qf_launch_count(self.h)
        f = np.zeros(max(n, 1))
        q = (ctypes.c_longlong * max(n, 1))()
        p = (ctypes.c_longlong * max(n, 1))()
        L.check(self.lib.qf_grad_launch_flops(self.h, L.dptr(f), q, p), "qf_grad_launch_flops")
        return f[:n], np.array(q[:n]), np.array(p[:n])

    def fock2e_os(self, Pa, Pb, cutmatrix=None):
        """Open-shell G for host densities (synchronous): (Ga, Gb) with
        Ga = J(Pa+Pb) - hyb K(Pa), Gb = J(Pa+Pb) - hyb K(Pb)."""
        self.prepare()
        Pa = np.ascontiguousarray(Pa, dtype=np.float64)
        Pb = np.ascontiguousarray(Pb, dtype=np.float64)
        Ga, Gb = np.zeros((self.n, self.n)), np.zeros((self.n, self.n))
        cm = None if cutmatrix is None else np.ascontiguousarray(cutmatrix, dtype=np.float64)
        cmp = L.dptr(cm) if cm is not None else ctypes.POINTER(ctypes.c_double)()
        L.check(self.lib.qf_fock_build_os_host(self.h, L.dptr(Pa), L.dptr(Pb), L.dptr(Ga), L.dptr(Gb), cmp),
                "qf_fock_build_os_host")
        return Ga, Gb

    def fock2e_os_device(self, Pa_ptr, Pb_ptr, Fa_ptr, Fb_ptr, stream=None, cutmat_ptr=None):
        """Fa += Ga, Fb += Gb with device pointers (ints); asynchronous."""
        self.prepare()
        v = ctypes.c_void_p
        L.check(self.lib.qf_fock_build_os_device(self.h, v(Pa_ptr), v(Pb_ptr), v(Fa_ptr), v(Fb_ptr),
                                                 v(cutmat_ptr or 0), v(stream or 0)), "qf_fock_build_os_device")

    def device_g_ptr(self):
        return self.lib.qf_device_g(self.h)

    def flops(self):
        """Algorithmic FLOPs of a build with the last density (per-class table)."""
        return self.lib.qf_flops_estimate(self.h)

    CLASS_NAMES = ["ss", "ps", "pp", "ds", "dp", "dd", "fs", "fp", "fd", "ff"]

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

    def set_tile(self, enable=True, waves=4, tasks_per_group=0, lds_bytes=64 * 1024):
        """LDS-tile digestion on/off; tasks_per_group 0: 4 up to 256 functions, 16 above."""
        L.check(self.lib.qf_set_tile(self.h, 1 if enable else 0, int(waves), int(tasks_per_group), int(lds_bytes)),
                "qf_set_tile")

    def set_class_tuning(self, cb, ck, tile=-1, waves=0, tasks_per_group=0, bras_per_task=0):
        L.check(self.lib.qf_set_class_tuning(self.h, int(cb), int(ck), int(tile), int(waves), int(tasks_per_group),
                                             int(bras_per_task)), "qf_set_class_tuning")

    def set_persistent(self, max_workgroups=0):
        L.check(self.lib.qf_set_persistent(self.h, int(max_workgroups)), "qf_set_persistent")

    def set_prim_split(self, piece=0):
        """Primitive pairs per task piece: 0 automatic, < 0 never split."""
        L.check(self.lib.qf_set_prim_split(self.h, int(piece)), "qf_set_prim_split")

    def set_class_piece(self, cb, ck, piece=0):
        """Primitive piece size of one class: 0 engine default, < 0 never split."""
        L.check(self.lib.qf_set_class_piece(self.h, int(cb), int(ck), int(piece)), "qf_set_class_piece")

    def set_class_os(self, cb, ck, mode=-1):
        """Open-shell form of class (cb|ck): 0 fused, 2/-1 fused lean (LDS-tile
        kernel: no per-task second-exchange state in registers), 1 two
        closed-shell passes."""
        L.check(self.lib.qf_set_class_os(self.h, int(cb), int(ck), int(mode)), "qf_set_class_os")

    def set_timing(self, on=True):
        L.check(self.lib.qf_set_timing(self.h, 1 if on else 0), "qf_set_timing")

    def read_timing(self):
        n = self.lib.qf_launch_count(self.h)
        ms = np.zeros(max(n, 1))
        calls = np.zeros(max(n, 1), dtype=np.int32)
        L.check(self.lib.qf_read_timing(self.h, L.dptr(ms), L.iptr(calls)), "qf_read_timing")
        return ms[:n], calls[:n]


class FockBatch:
    """Many molecules built together: F_i += G(P_i) for every engine with one
    launch per quartet class (``qf_batch_*``), optionally replayed from a HIP
    graph.  Densities and Fock accumulators live in batch-owned device slots
    (``density_ptr(i)``, ``fock_ptr(i)``); ``build(stream)`` is asynchronous.
    """

    def __init__(self, engines, replicas=2, graph=False, density_ptrs=None, fock_ptrs=None):
        self.lib = L.lib()
        self.engines = list(engines)            # keeps the engines alive
        m = len(self.engines)
        arr = (ctypes.c_void_p * m)(*[e.h for e in self.engines])
        h = ctypes.c_void_p()
        if density_ptrs is None:
            L.check(self.lib.qf_batch_create(arr, m, int(replicas), ctypes.byref(h)), "qf_batch_create")
        else:
            dp = (ctypes.c_void_p * m)(*density_ptrs)
            fp = (ctypes.c_void_p * m)(*fock_ptrs)
            L.check(self.lib.qf_batch_create_ext(arr, m, int(replicas), dp, fp, ctypes.byref(h)),
                    "qf_batch_create_ext")
        self.h = h
        self.n = [e.n for e in self.engines]
        self.set_graph(graph)

    def close(self):
        if getattr(self, "h", None):
            self.lib.qf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.engines)

    def set_graph(self, on=True):
        L.check(self.lib.qf_batch_set_graph(self.h, 1 if on else 0), "qf_batch_set_graph")

    def density_ptr(self, i):
        return self.lib.qf_batch_density(self.h, i)

    def fock_ptr(self, i):
        return self.lib.qf_batch_fock(self.h, i)

    def g_ptr(self, i):
        return self.lib.qf_batch_g(self.h, i)

    def build(self, stream=None):
        L.check(self.lib.qf_batch_build(self.h, ctypes.c_void_p(stream or 0)), "qf_batch_build")

    def set_density(self, i, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == (self.n[i], self.n[i])
        L.check(self.lib.qf_batch_set_density_host(self.h, i, L.dptr(P)), "qf_batch_set_density_host")

    def get(self, i, g=True, f=False):
        """(G_i, F_i) as host arrays (synchronous)."""
        n = self.n[i]
        G = np.zeros((n, n)) if g else None
        F = np.zeros((n, n)) if f else None
        null = ctypes.POINTER(ctypes.c_double)()
        L.check(self.lib.qf_batch_get_host(self.h, i, L.dptr(G) if g else null, L.dptr(F) if f else null),
                "qf_batch_get_host")
        return G, F

    def clear_fock(self, stream=None):
        L.check(self.lib.qf_batch_clear_fock(self.h, ctypes.c_void_p(stream or 0)), "qf_batch_clear_fock")

    def launches(self):
        out = []
        for k in range(self.lib.qf_batch_launch_count(self.h)):
            cb, ck, f32, nt = (ctypes.c_int() for _ in range(4))
            L.check(self.lib.qf_batch_launch_info(self.h, k, ctypes.byref(cb), ctypes.byref(ck), ctypes.byref(f32),
                                                  ctypes.byref(nt)), "qf_batch_launch_info")
            out.append(("(%s|%s)" % (FockEngine.CLASS_NAMES[cb.value], FockEngine.CLASS_NAMES[ck.value]),
                        bool(f32.value), nt.value))
        return out

    def set_timing(self, on=True):
        L.check(self.lib.qf_batch_set_timing(self.h, 1 if on else 0), "qf_batch_set_timing")

    def read_timing(self):
        n = self.lib.qf_batch_launch_count(self.h)
        ms = np.zeros(max(n, 1))
        calls = np.zeros(max(n, 1), dtype=np.int32)
        L.check(self.lib.qf_batch_read_timing(self.h, L.dptr(ms), L.iptr(calls)), "qf_batch_read_timing")
        return ms[:n], calls[:n]
