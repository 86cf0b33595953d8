"""CPU-side checks of the product: the C ABI library loads and exports what
include/quickfock.h declares, the host basis layout equals QUICK's, the
geometry generators are deterministic, and the engine refuses to run
without a GPU (no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from quick_amd import basis as B
from quick_amd import molecules as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "quick_amd", "libquickfock.so")
HDR = os.path.join(ROOT, "include", "quickfock.h")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", txt)
    return sorted(set(n for n in names if n.startswith(("qf_", "gpu_"))))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with python -m quick_amd.build"
    from quick_amd import _lib
    assert os.path.samefile(_lib.LIBPATH, LIB)
    _lib.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    decl = declared_functions()
    assert len(decl) >= 30
    missing = [d for d in decl if d not in exported]
    assert not missing, missing


def test_engine_fails_loudly_without_gpu():
    from quick_amd import _lib
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    rc = _lib.lib().qf_engine_create(0, ctypes.byref(h))
    assert rc != 0


@pytest.mark.parametrize("basis", ["STO-3G", "3-21G", "6-31G", "6-31G*", "6-31G**", "cc-pVDZ", "def2-SVP"])
def test_product_basis_equals_quick_layout(basis):
    """quick_amd.basis (the product's own normalisation) against the oracle's
    literal restatement of QUICK's reader (xnorm/xnewnorm, gccoeff(ksumtype+L))."""
    mol = M.polyglycine(1) if basis != "def2-SVP" else M.alkane(2)
    bs = B.build_basis(mol, basis)
    qb = O.QuickBasisRef(mol.symbols, mol.xyz, basis)
    a = bs.shell_arrays()
    assert bs.nbasis == qb.nbasis and len(bs.shells) == qb.nshell and bs.nprim == qb.jbasis
    assert np.array_equal(a["atom"], qb.katom)
    assert np.array_equal(a["lmin"], qb.qstart) and np.array_equal(a["lmax"], qb.qfinal)
    assert np.array_equal(a["pstart"], qb.kstart) and np.array_equal(a["fstart"], qb.ksumtype)
    assert np.array_equal(a["exps"], qb.gcexpo)
    assert np.allclose(a["coef"], qb.gccoeff, rtol=1e-14, atol=0)


def test_generators_are_deterministic_and_sane():
    a, b = M.water_cluster(32), M.water_cluster(32)
    assert np.array_equal(a.xyz, b.xyz)
    o = a.xyz[0::3] * B.BOHR
    d = np.linalg.norm(o[:, None] - o[None], axis=2) + np.eye(32) * 10
    assert d.min() >= 2.7 - 1e-9
    bs = B.build_basis(a, "6-31G*")
    assert bs.nbasis == 608 and a.natom == 96
    g = M.polyglycine(10)
    from collections import Counter
    assert Counter(g.symbols) == Counter({"C": 20, "H": 32, "N": 10, "O": 11})
    xyz = g.xyz * B.BOHR
    dd = np.linalg.norm(xyz[:, None] - xyz[None], axis=2) + np.eye(len(xyz)) * 10
    assert dd.min() > 0.9
    for n in range(1, 13):
        m = M.alkane(n)
        assert Counter(m.symbols) == Counter({"C": n, "H": 2 * n + 2})
        x = m.xyz * B.BOHR
        dd = np.linalg.norm(x[:, None] - x[None], axis=2) + np.eye(len(x)) * 10
        assert dd.min() > 1.0


def test_codegen_flop_table_is_committed_and_consistent():
    import json
    from quick_amd.codegen import gen_eri
    tab = json.load(open(os.path.join(ROOT, "quick_amd", "csrc", "gen", "flop_table.json")))
    grad = tab.pop("gradient")
    assert len(tab) == 55          # (La Lb|Lc Ld) for s, p, d, f
    assert len(grad) == 21         # gradient classes: s, p, d
    from quick_amd.codegen import gen_grad
    for name, info in grad.items():
        t = gen_grad.grad_tables(*(int(ch) for ch in name[1:]))
        assert (t["prim_flop"], t["quartet_flop"]) == (info["prim_flop"], info["quartet_flop"])
    for name, info in tab.items():
        La, Lb, Lc, Ld = info["La"], info["Lb"], info["Lc"], info["Ld"]
        _, _, fresh = gen_eri.gen_class(La, Lb, Lc, Ld)
        assert fresh["vrr_flop"] == info["vrr_flop"] and fresh["hrr_flop"] == info["hrr_flop"]
        nc = lambda L: (L + 1) * (L + 2) // 2
        assert info["ni"] == nc(La) * nc(Lb) * nc(Lc) * nc(Ld)


def test_autotune_table_round_trip(tmp_path):
    """A tuned entry is stored under its key, other keys survive, and the
    loaded entry is what was stored."""
    from quick_amd import autotune as A
    path = str(tmp_path / "tab.json")
    chosen = {"(ss|ss)": ((1, 4, 8, 0, 16), 1.5), "(dd|dd)": ((2, 0, 0, 1, 8), 0.2)}
    A.save_table("dev|single|fp64|n>256|x", chosen, path)
    A.save_table("dev|batch|mixed|n<=128", {"(ps|ss)": ((0, 0, 0, 2, -1), 0.1)}, path)
    got = A.load_table("dev|single|fp64|n>256|x", path)
    assert got == {"(ss|ss)": (1, 4, 8, 0, 16), "(dd|dd)": (2, 0, 0, 1, 8)}
    assert A.load_table("dev|batch|mixed|n<=128", path) == {"(ps|ss)": (0, 0, 0, 2, -1)}
    assert A.load_table("missing", path) is None
    assert A.size_bucket(128) == "n<=128" and A.size_bucket(200) == "n<=256" and A.size_bucket(608) == "n>256"


class _MockEngine:
    """Stands in for a FockEngine in the autotuner: per-class launch times
    follow a fixed cost table of the current settings (no GPU)."""
    NAMES = {(0, 0): "(ss|ss)", (1, 0): "(ps|ss)"}

    def __init__(self):
        self.tune = {k: (-1, 0, 0, 0) for k in self.NAMES}
        self.piece = {k: 0 for k in self.NAMES}
        self.os = {k: -1 for k in self.NAMES}
        self.split = {k: -1 for k in self.NAMES}
        self.acc = None

    def set_class_tuning(self, cb, ck, tile=-1, waves=0, tpg=0, mmax=0):
        if (cb, ck) in self.NAMES:
            self.tune[(cb, ck)] = (tile, waves, tpg, mmax)

    def set_class_piece(self, cb, ck, piece):
        if (cb, ck) in self.NAMES:
            self.piece[(cb, ck)] = piece

    def set_class_os(self, cb, ck, mode):
        if (cb, ck) in self.NAMES:
            self.os[(cb, ck)] = mode

    def set_class_split(self, cb, ck, mode):
        if (cb, ck) in self.NAMES:
            self.split[(cb, ck)] = mode

    def launches(self):
        out = []
        for k, nm in self.NAMES.items():
            out.append((nm, False, 1))
            if self.split[k] == 1:
                out.append((nm, True, 1))
        return out

    def cost(self, k, fp32):
        # (ss|ss) is fastest on the tile kernel with pieces of 8, fused lean
        # and unsplit; (ps|ss) on the wave-per-quartet kernel, unsplit pieces,
        # two passes and split (its two halves cost less than one FP64 launch)
        best_tile, best_piece, best_os, split_gain = {(0, 0): (1, 8, 2, -0.1), (1, 0): (2, -1, 1, 0.3)}[k]
        t = 1.0 + (0.0 if self.tune[k][0] == best_tile else 0.5) + (0.0 if self.piece[k] == best_piece else 0.2)
        t += 0.0 if self.os[k] in (-1, best_os) else 0.1
        if self.split[k] == 1:
            return 0.5 * (t - split_gain)   # each half
        return t

    def set_timing(self, on=True):
        self.acc = [[0.0, 0] for _ in self.launches()] if on else None

    def build(self):
        if self.acc is not None:
            for i, (nm, f32, _) in enumerate(self.launches()):
                k = next(k for k, v in self.NAMES.items() if v == nm)
                self.acc[i][0] += self.cost(k, f32)
                self.acc[i][1] += 1

    def read_timing(self):
        import numpy as np
        return np.array([a[0] for a in self.acc]), np.array([a[1] for a in self.acc])


def test_autotune_passes_choose_per_class_and_table_applies_them(tmp_path):
    """autotune(): the shape, piece, open-shell and split passes each keep the
    fastest setting per class (split timed over both precision halves); the
    candidate tuple carries them in that order and apply_table sets them."""
    from quick_amd import autotune as A
    e = _MockEngine()
    cands = [(1, 4, 8, 0), (2, 0, 0, 1)]
    chosen = A.autotune(e, 0, builds=2, candidates=cands, pieces=[0, 8, -1], build=e.build,
                        os_modes=(0, 1, 2), split_modes=(0, 1))
    assert chosen["(ss|ss)"][0] == (1, 4, 8, 0, 8, 2, 0)
    assert chosen["(ps|ss)"][0] == (2, 0, 0, 1, -1, 1, 1)
    path = str(tmp_path / "tab.json")
    A.save_table("k", chosen, path)
    f = _MockEngine()
    A.apply_table(f, A.load_table("k", path))
    assert f.tune[(0, 0)] == (1, 4, 8, 0) and f.piece[(0, 0)] == 8 and f.os[(0, 0)] == 2 and f.split[(0, 0)] == 0
    assert f.tune[(1, 0)] == (2, 0, 0, 1) and f.piece[(1, 0)] == -1 and f.os[(1, 0)] == 1 and f.split[(1, 0)] == 1
    # without an open-shell pass the split mode follows a -1 placeholder
    g = _MockEngine()
    chosen = A.autotune(g, 0, builds=1, candidates=cands, pieces=[0, 8, -1], build=g.build, split_modes=(0, 1))
    assert chosen["(ps|ss)"][0] == (2, 0, 0, 1, -1, -1, 1)


def test_tune_grad_keeps_the_faster_kernel_per_class():
    """tune_grad(): per class, the gradient kernel (per-lane 1, wave-per-quartet
    0) with the smaller event-timed sum is left set."""
    from quick_amd import autotune as A

    class G(_MockEngine):
        def __init__(self):
            super().__init__()
            self.grad = {k: -1 for k in self.NAMES}

        def set_grad_class(self, cb, ck, mode):
            self.grad[(cb, ck)] = mode

        def cost(self, k, fp32):
            # (ss|ss) faster per lane, (ps|ss) on the wave-per-quartet kernel
            return {(0, 0): {1: 1.0, 0: 3.0}, (1, 0): {1: 2.0, 0: 1.5}}[k][self.grad[k]]

    e = G()
    chosen = A.tune_grad(e, e.build, calls=2)
    assert chosen["(ss|ss)"][0] == 1 and chosen["(ps|ss)"][0] == 0
    assert e.grad[(0, 0)] == 1 and e.grad[(1, 0)] == 0
