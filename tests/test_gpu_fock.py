"""Parity of the HIP Fock build (through the C ABI) with the CPU oracle.

Bar: FP64 Fock matrix max |dF| <= 1e-10 Eh against QUICK's CPU get2e
restated (same geometry, basis, density, cutoffs); SCF energy <= 1e-8 Eh.
"""
import ctypes

import numpy as np
import pytest

from tests.helpers import announce, molecule, oracle_for, product_basis, quick_abi_setup, seeded_density_for

pytestmark = pytest.mark.gpu

FOCK_TOL = 1e-10


def _engine(bs, **kw):
    from quick_amd.engine import FockEngine
    return FockEngine(bs, **kw)


CASES = [
    ("water", "6-31G*", 1e-12),
    ("water", "cc-pVDZ", 1e-12),
    ("water", "STO-3G", 1e-12),
    ("water3", "6-31G*", 1e-12),
    ("alkane2", "6-31G**", 1e-10),
    ("gly1", "6-31G*", 1e-11),
    ("LiF", "3-21G", 1e-12),
    ("NH4+", "def2-SVP", 1e-12),
]


@pytest.mark.parametrize("name,basis,cutoff", CASES)
def test_schwarz_bounds_match_oracle(name, basis, cutoff):
    announce(f"schwarz {name} {basis}")
    mol = molecule(name)
    orc = oracle_for(mol, basis)
    eng = _engine(product_basis(mol, basis), cutoff=cutoff)
    y_gpu, c_gpu = eng.schwarz()
    y_ref, c_ref = orc.schwarz()
    assert np.allclose(y_gpu, y_ref, rtol=1e-11, atol=1e-14)
    mask = c_ref > 0          # QUICK fills cutprim for shell order II <= JJ only
    assert np.allclose(c_gpu[mask], c_ref[mask], rtol=1e-11, atol=1e-14)
    assert np.allclose(c_gpu, c_gpu.T)


@pytest.mark.parametrize("name,basis,cutoff", CASES)
def test_fock_matches_oracle(name, basis, cutoff):
    announce(f"fock {name} {basis} cutoff={cutoff}")
    mol = molecule(name)
    orc = oracle_for(mol, basis)
    P, _ = seeded_density_for(orc, mol)
    G_ref, nq = orc.fock2e(P, cutoff=cutoff, threads=4)
    eng = _engine(product_basis(mol, basis), cutoff=cutoff)
    G = eng.fock2e(P)
    err = float(np.max(np.abs(G - G_ref)))
    print(f"  max|dF| = {err:.3e}  (quartets {nq})", flush=True)
    assert err <= FOCK_TOL
    assert np.max(np.abs(G - G.T)) <= 1e-14 * max(1.0, np.max(np.abs(G)))


@pytest.mark.parametrize("mode,name", [("direct", "water3"), ("tile-small", "water3"), ("tile-1wave", "water3"),
                                       ("coop", "water3"), ("coop", "gly1"), ("coop-4bras", "alkane2")])
def test_digestion_paths_match_oracle(mode, name):
    """The direct-atomic kernels, LDS tiles forced small (some classes fall
    back, chunks shrink), one-wave tile workgroups and the wave-per-quartet
    kernels (d-heavy classes; the rest on tiles) all reproduce the oracle on
    molecules with s, p and d classes."""
    announce(f"digestion path {mode} {name} 6-31G*")
    mol = molecule(name)
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-12, threads=4)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-12)
    if mode == "direct":
        eng.set_tile(False)
    elif mode == "tile-small":
        eng.set_tile(True, waves=2, tasks_per_group=3, lds_bytes=4096)
    elif mode == "tile-1wave":
        eng.set_tile(True, waves=1, tasks_per_group=1)
    else:
        for cb in range(6):
            for ck in range(cb + 1):
                eng.set_class_tuning(cb, ck, 2, 0, 0, 4 if mode == "coop-4bras" else 0)
    G = eng.fock2e(P)
    err = float(np.max(np.abs(G - G_ref)))
    print(f"  max|dF| = {err:.3e}", flush=True)
    assert err <= FOCK_TOL


@pytest.mark.parametrize("mode", ["default", "direct", "coop"])
@pytest.mark.parametrize("piece", [1, 2, 6])
def test_primitive_pieces_match_oracle(piece, mode):
    """Quartets whose primitive pairs are spread over several tasks (pieces of
    1, 2 or 6 bra and ket primitive pairs) give the same Fock matrix as one
    task per quartet and as the oracle: the HRR and the digestion are linear
    in the primitive sums."""
    announce(f"primitive pieces {piece} {mode} water3 6-31G*")
    mol = molecule("water3")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-12, threads=4)
    bs = product_basis(mol, "6-31G*")

    def engine(split):
        e = _engine(bs, cutoff=1e-12)
        e.set_prim_split(split)
        if mode == "direct":
            e.set_tile(False)
        elif mode == "coop":
            for cb in range(6):
                for ck in range(cb + 1):
                    e.set_class_tuning(cb, ck, 2, 0, 0, 0)
        return e
    eng = engine(-1)            # whole quartets, same kernels
    G_whole = eng.fock2e(P)
    eng2 = engine(piece)
    G = eng2.fock2e(P)
    assert eng2.stats()["n_tasks_fp64"] > eng.stats()["n_tasks_fp64"]
    err, dw = float(np.max(np.abs(G - G_ref))), float(np.max(np.abs(G - G_whole)))
    print(f"  max|dF| vs oracle {err:.3e}, vs whole quartets {dw:.3e}", flush=True)
    assert err <= FOCK_TOL and dw <= 1e-12


def test_config1_water_fock_with_quick_schwarz_arrays():
    """Drop-in path: Ycutoff/cutprim as QUICK computes them (one triangle of
    cutprim filled) handed to the engine instead of the GPU's own."""
    announce("config1 water 6-31G* with oracle Schwarz arrays")
    mol = molecule("water")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-12)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-12, schwarz=orc.schwarz())
    G = eng.fock2e(P)
    assert np.max(np.abs(G - G_ref)) <= FOCK_TOL


def test_one_electron_matrices_match_oracle():
    announce("one-electron S T V water 6-31G* and def2-SVP NH4+")
    for name, basis in [("water", "6-31G*"), ("NH4+", "def2-SVP"), ("gly1", "6-31G**")]:
        mol = molecule(name)
        orc = oracle_for(mol, basis)
        eng = _engine(product_basis(mol, basis))
        for a, b in zip(eng.one_electron(), orc.one_electron()):
            assert np.max(np.abs(a - b)) < 1e-11


def test_fock_is_linear_in_density_and_symmetric():
    """Size-independent properties on the 16-water cluster: G(P1) + G(P2-P1)
    = G(P2) (the incremental build QUICK uses after ncyc) and symmetry.
    Screening uses the density, so compare with screening off."""
    announce("linearity water16 6-31G*")
    mol = molecule("water16")
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-30, prim_limit=1e-31)
    S, _, _ = eng.one_electron()
    from quick_amd.scf import seeded_density
    P1 = seeded_density(S, mol.nelec() // 2, seed=1)
    P2 = seeded_density(S, mol.nelec() // 2, seed=2)
    G1, G2, G12 = eng.fock2e(P1), eng.fock2e(P2), eng.fock2e(P2 - P1)
    scale = np.max(np.abs(G2))
    assert np.max(np.abs(G1 + G12 - G2)) < 1e-12 * scale
    assert np.max(np.abs(G2 - G2.T)) < 1e-14 * scale


def test_scf_energy_matches_oracle_config1():
    """Converged RHF energy, GPU Fock builds vs oracle Fock builds, same
    cutoff; and QUICK's saved value for its own H2O/6-31G* test geometry."""
    announce("scf water 6-31G*")
    from oracle import oracle as O
    from quick_amd import basis as B
    from quick_amd.scf import rhf
    xyz = np.array([[-0.33840, 0.00380, 0.23923], [-0.33510, -0.00190, -0.83277], [0.67350, -0.00190, 0.59353]])
    mol = B.Molecule.from_angstrom(["O", "H", "H"], xyz)
    orc = oracle_for(mol, "6-31G*")
    E_ref, _, _ = O.rhf(orc, 10, cutoff=1e-10, prim_limit=1e-10, etol=1e-11, dtol=1e-9)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-10, prim_limit=1e-10)
    res = rhf(eng, etol=1e-11, dtol=1e-9)
    assert res.converged
    print(f"  E_gpu {res.energy:.12f}  E_oracle {E_ref:.12f}", flush=True)
    assert abs(res.energy - E_ref) <= 1e-8
    assert abs(res.energy - (-75.983996360)) < 3e-8


@pytest.mark.parametrize("precision,tol", [("fp64", 2e-8), ("mixed", 1e-6)])
def test_device_scf_matches_host_scf(precision, tol):
    """Device-resident SCF with incremental Fock builds (QUICK's deltaO)
    against the host SCF with full FP64 builds, water octamer 6-31G*."""
    announce(f"device SCF ({precision}, incremental) vs host SCF, water8")
    from quick_amd.scf import rhf, rhf_device
    mol = molecule("water8")
    bs = product_basis(mol, "6-31G*")
    e64 = _engine(bs, cutoff=1e-11, prim_limit=1e-12)
    S, T, V = e64.one_electron()
    ref = rhf(e64, etol=1e-9, dtol=1e-7, one_electron=(S, T, V))
    e = _engine(bs, cutoff=1e-11, prim_limit=1e-12, precision=precision)
    res = rhf_device(e, etol=1e-9, dtol=1e-7, one_electron=(S, T, V))
    print(f"  host {ref.energy:.10f} ({ref.iterations} it)  device {res.energy:.10f} ({res.iterations} it, "
          f"{res.full_builds} full builds)  dE {abs(res.energy - ref.energy):.1e}", flush=True)
    assert res.converged and ref.converged
    assert abs(res.energy - ref.energy) <= tol
    assert res.full_builds < res.iterations


def test_mixed_precision_error_is_bounded():
    """Mixed precision (FP32 VRR/HRR below 1e-4 of the largest Schwarz
    product) against the FP64 oracle, water octamer 6-31G*: the error of
    the Fock matrix and of the two-electron energy in Eh."""
    announce("mixed precision water8 6-31G* vs oracle")
    mol = molecule("water8")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-12, threads=8)
    emx = _engine(product_basis(mol, "6-31G*"), cutoff=1e-12, precision="mixed", split=1e-4)
    Gmx = emx.fock2e(P)
    st = emx.stats()
    assert st["n_tasks_fp32"] > 0
    dF = float(np.max(np.abs(Gmx - G_ref)))
    dE = 0.5 * abs(float(np.sum(P * (Gmx - G_ref))))
    print(f"  mixed vs oracle: max|dF| {dF:.2e}  |dE2| {dE:.2e} Eh", flush=True)
    assert dF < 1e-6
    assert dE < 1e-6


@pytest.mark.parametrize("name", ["water3", "alkane3"])
def test_mixed_precision_with_direct_fp64_launches(name):
    """Tile digestion off: the FP64 halves run through the direct-atomic
    kernels while every FP32 half still runs through the LDS-tile kernels;
    checked against the FP64 oracle."""
    announce(f"mixed precision, direct FP64 launches, {name} 6-31G* vs oracle")
    mol = molecule(name)
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-10, threads=4)
    emx = _engine(product_basis(mol, "6-31G*"), cutoff=1e-10, precision="mixed")
    emx.set_tile(False)
    G = emx.fock2e(P)
    assert emx.stats()["n_tasks_fp32"] > 0
    dF = float(np.max(np.abs(G - G_ref)))
    dE = 0.5 * abs(float(np.sum(P * (G - G_ref))))
    print(f"  max|dF| {dF:.2e}  |dE2| {dE:.2e} Eh", flush=True)
    assert dF < 1e-6 and dE < 1e-6


@pytest.mark.skip(reason="the direct-atomic FP32 kernels are not built: the (dp|dp) one gave max|dF| of "
                         "1.6-3.0 Eh on gfx950 and the (dd|dp) one non-finite values, while the FP64 direct "
                         "and FP32 LDS-tile kernels of the same classes were right; the cause was not found, so "
                         "every FP32 half runs on the LDS-tile kernel and a direct FP32 launch is refused")
def test_direct_fp32_kernels_match_oracle():
    pass


def test_scf_energy_matches_oracle_on_water_cluster():
    """The water-cluster generator of the 32-water workload at a size the
    oracle converges in seconds: GPU SCF (FP64 builds) against the oracle's
    SCF, both with full builds at cutoff 1e-10, to <= 1e-8 Eh."""
    announce("scf (H2O)4 6-31G* GPU vs oracle")
    from oracle import oracle as O
    from quick_amd.scf import rhf
    mol = molecule("water4")
    orc = oracle_for(mol, "6-31G*")
    E_ref, _, _ = O.rhf(orc, mol.nelec(), cutoff=1e-10, prim_limit=1e-11, etol=1e-11, dtol=1e-9, threads=8)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-10, prim_limit=1e-11)
    res = rhf(eng, etol=1e-11, dtol=1e-9)
    assert res.converged
    print(f"  E_gpu {res.energy:.12f}  E_oracle {E_ref:.12f}  dE {abs(res.energy - E_ref):.1e}", flush=True)
    assert abs(res.energy - E_ref) <= 1e-8


def test_quick_fortran_abi_drop_in():
    """Call the gpu_*_ entry points the way QUICK's Fortran does: 1-based
    shell arrays, gccoeff/gcexpo(MAXPRIM=14, nbasis) column-major,
    YCutoff/cutPrim from QUICK's schwarzoff, o += G(dense)."""
    announce("QUICK Fortran ABI water 6-31G*")
    from quick_amd import _lib
    lib = _lib.lib()
    mol = molecule("water")
    orc = oracle_for(mol, "6-31G*")
    qb = orc.qb
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-12)
    n = qb.nbasis
    vp, D = quick_abi_setup(lib, mol, orc)
    dummy_d = np.zeros(max(n, 1) * 4, dtype=np.float64)
    f = ctypes.c_bool(False)
    ierr = ctypes.c_int(0)
    o = np.asfortranarray(np.full((n, n), 0.25))
    dense = np.asfortranarray(P)
    lib.gpu_upload_calculated_(vp(o), vp(dummy_d), vp(dummy_d), dense.ctypes.data_as(ctypes.c_void_p))
    cm = np.asfortranarray(orc.density_cutoff(P))
    lib.gpu_upload_cutoff_(cm.ctypes.data_as(ctypes.c_void_p), D(1e-12), D(1e-13), D(1e-10), D(1e-12))
    lib.gpu_get_cshell_eri_(ctypes.byref(f), o.ctypes.data_as(ctypes.c_void_p))
    lib.gpu_delete_(ctypes.byref(ierr))
    assert np.max(np.abs((o - 0.25) - G_ref)) <= FOCK_TOL


def test_quick_oei_and_device_info_entry_points():
    """gpu_get_oei_ adds the nuclear attraction of the atoms and of external
    point charges into o (QUICK's gpu_get2e_getxc_drivers.h:360); an
    external charge q on top of atom C must equal raising Z_C by q.
    gpu_get_device_info_ fills QUICK's fields."""
    announce("QUICK OEI / device-info entry points, water 6-31G*")
    from quick_amd import _lib
    lib = _lib.lib()
    mol = molecule("water")
    orc = oracle_for(mol, "6-31G*")
    S_ref, T_ref, V_ref = orc.one_electron()
    qb = orc.qb
    n = qb.nbasis
    I = lambda v: ctypes.byref(ctypes.c_int(int(v)))
    ierr = ctypes.c_int(0)
    lib.gpu_new_(ctypes.byref(ierr))
    lib.gpu_set_device_(I(0), ctypes.byref(ierr))
    lib.gpu_init_device_(ctypes.byref(ierr))
    assert ierr.value == 0
    # device info (Fortran character(len=256) buffers)
    cnt, dev, mem, nproc, nl, al, mj, mn = (ctypes.c_int() for _ in range(8))
    freq = ctypes.c_double()
    name, arch = ctypes.create_string_buffer(256), ctypes.create_string_buffer(256)
    lib.gpu_get_device_info_(ctypes.byref(cnt), ctypes.byref(dev), ctypes.byref(mem), ctypes.byref(nproc),
                             ctypes.byref(freq), name, ctypes.byref(nl), arch, ctypes.byref(al), ctypes.byref(mj),
                             ctypes.byref(mn), ctypes.byref(ierr))
    assert ierr.value == 0 and cnt.value >= 1 and nproc.value > 0 and mem.value > 0
    assert "gfx950" in arch.raw[:al.value].decode()
    print(f"  {name.raw[:nl.value].decode()} {arch.raw[:al.value].decode()} CUs {nproc.value}", flush=True)
    # basis upload through the QUICK shim, then o += V
    lib.gpu_setup_(I(len(mol.symbols)), I(n), I(10), I(1), I(0), I(2))
    xyz = np.ascontiguousarray(qb.xyz)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.gpu_upload_xyz_(vp(xyz))
    chg = np.ascontiguousarray(qb.charge, dtype=np.float64)
    atoms = np.zeros(len(mol.symbols), dtype=np.int32)
    lib.gpu_upload_atom_and_chg_(vp(atoms), vp(chg))
    MAXPRIM = 14
    gccoeff = np.zeros((n, MAXPRIM))
    gcexpo = np.zeros((n, MAXPRIM))
    for s in range(qb.nshell):
        f0 = qb.ksumtype[s]
        nf = sum((L + 1) * (L + 2) // 2 for L in range(qb.qstart[s], qb.qfinal[s] + 1))
        p0, k = qb.kstart[s], qb.kprim[s]
        for fi in range(f0, f0 + nf):
            gcexpo[fi, :k] = qb.gcexpo[p0:p0 + k]
        for L in range(qb.qstart[s], qb.qfinal[s] + 1):
            gccoeff[f0 + L, :k] = qb.gccoeff[p0:p0 + k, L]
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    dummy_i = np.zeros(max(n, qb.nshell) * 4 + 8, dtype=np.int32)
    dummy_d = np.zeros(max(n, 1) * 4, dtype=np.float64)
    lib.gpu_upload_basis_.restype = None
    lib.gpu_upload_basis_(I(qb.nshell), I(qb.jbasis), I(qb.nshell), I(qb.jbasis), I(int(qb.kprim.max())),
                          vp(dummy_i), vp(dummy_i), vp(dummy_d), vp(dummy_d), vp(dummy_i), vp(dummy_i), vp(dummy_i),
                          vp(dummy_i), vp(dummy_i), vp(i32(qb.kstart + 1)), vp(i32(qb.katom + 1)), vp(dummy_i),
                          vp(i32(qb.kprim)), vp(dummy_i), vp(i32(qb.ksumtype + 1)), vp(dummy_i), vp(i32(qb.qstart)),
                          vp(i32(qb.qfinal)), vp(dummy_i), vp(dummy_i), vp(gccoeff), vp(dummy_d), vp(gcexpo),
                          vp(dummy_i))
    o = np.asfortranarray(np.full((n, n), 0.5))
    lib.gpu_upload_oei_(I(0), vp(dummy_d), vp(dummy_d), ctypes.byref(ierr))
    lib.gpu_get_oei_(vp(o))
    assert np.max(np.abs((o - 0.5) - V_ref)) <= 1e-11
    # an external charge of 0.75 on the oxygen == Z_O + 0.75
    ext = np.ascontiguousarray(qb.xyz[0].copy())
    q = np.array([0.75])
    lib.gpu_upload_oei_(I(1), vp(ext), vp(q), ctypes.byref(ierr))
    o2 = np.asfortranarray(np.zeros((n, n)))
    lib.gpu_get_oei_(vp(o2))
    chg2 = chg.copy()
    chg2[0] += 0.75
    lib.gpu_upload_oei_(I(0), vp(dummy_d), vp(dummy_d), ctypes.byref(ierr))
    lib.gpu_upload_atom_and_chg_(vp(atoms), vp(chg2))
    o3 = np.asfortranarray(np.zeros((n, n)))
    lib.gpu_get_oei_(vp(o3))
    assert np.max(np.abs(o2 - o3)) <= 1e-12
    lib.gpu_cleanup_()
    lib.gpu_delete_(ctypes.byref(ierr))
