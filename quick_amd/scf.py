"""Restricted Hartree-Fock around the GPU Fock build.

Restates the control flow of QUICK's ``electdiis`` (quick_scf_module.f90:
344-760) and ``scf_operator`` (quick_scf_operator_module.f90:26-250):

* F = H + G(P); after ``ncyc`` full builds optionally incremental,
  F = F_prev + G(P - P_prev) (``deltaO``, quick_scf_operator_module.f90:69-79),
* E = 1/2 sum P (H + F) + E_nuc (eri.fh:1064-1083),
* Pulay DIIS on the orthogonalised commutator X^T (F P S - S P F) X with up to
  ``ndiis`` vectors, falling back to the plain Fock matrix when the DIIS
  system is singular (quick_scf_module.f90:405-602),
* X = S^-1/2, P = 2 C_occ C_occ^T.

``uhf`` is the unrestricted counterpart (QUICK's uscf / uscf_operator,
quick_uscf_operator_module.f90:26-250): F_a = H + J(P_a + P_b) - K(P_a),
F_b likewise, one DIIS over the stacked alpha and beta error vectors, and
E = 1/2 sum_s P_s (H + F_s) + E_nuc (get_oshell_eri_energy, eri.fh:1062-1080).

The Fock build runs on the GPU through :class:`quick_amd.engine.FockEngine`;
diagonalisation and DIIS (O(n^3), FP64) use LAPACK on the host.
"""
from __future__ import annotations

import time

import numpy as np


# QUICK's SCF defaults (quick_method_module.f90:96-126): level shifting from
# DIIS cycle 3 when the DIIS error exceeds 0.1 and the HOMO-LUMO gap is below
# 0.2 Eh; LSOLVE's pivot threshold is DMCutoff.
LSHIFT_CYCLE, LSHIFT_ERR, LSHIFT_GAP = 3, 0.1, 0.2
DM_CUTOFF = 1e-10


def lsolve(A, b, thresh=DM_CUTOFF):
    """QUICK's LSOLVE (Lsolve.f90:13-145): Gaussian elimination with row
    pivoting; None when a pivot falls below thresh times the mean |A_ij|."""
    A = np.array(A, dtype=np.float64)
    b = np.array(b, dtype=np.float64)
    n = len(b)
    if n == 1:
        return None if abs(A[0, 0]) < thresh else b / A[0, 0]
    ave = float(np.mean(np.abs(A)))
    if ave == 0.0:
        return None
    amin = ave * thresh
    for k in range(n - 1):
        irow = k + int(np.argmax(np.abs(A[k:, k])))
        if abs(A[irow, k]) < amin:
            return None
        if irow != k:
            A[[k, irow]] = A[[irow, k]]
            b[[k, irow]] = b[[irow, k]]
        t = A[k + 1:, k] / A[k, k]
        A[k + 1:, k + 1:] -= np.outer(t, A[k, k + 1:])
        b[k + 1:] -= t * b[k]
        A[k + 1:, k] = 0.0
    if abs(A[n - 1, n - 1]) < amin:
        return None
    x = np.zeros(n)
    for i in range(n - 1, -1, -1):
        x[i] = (b[i] - A[i, i + 1:] @ x[i + 1:]) / A[i, i]
    return x


def diis_coefficients(B, thresh=DM_CUTOFF):
    """Pulay coefficients for the k x k error-overlap matrix B
    (quick_scf_module.f90:540-586): solve [B -1; -1 0] c = [0 -1]; when LSOLVE
    fails, drop the oldest error vector and retry.  Returns (first, c): the
    coefficients apply to the vectors first..k-1; c is None when every
    reduced system failed (the caller keeps the plain Fock matrix)."""
    k = B.shape[0]
    for first in range(k):
        m = k - first
        A = -np.ones((m + 1, m + 1))
        A[:m, :m] = B[first:, first:]
        A[m, m] = 0.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        c = lsolve(A, rhs, thresh)
        if c is not None:
            return first, c[:m]
    return k, None


def adjust_cutoff(prms, cutoff, prim_limit):
    """QUICK's adjust_Cutoff (quick_method_module.f90:1105-1128): tighten the
    integral cutoff to 1e-9, 1e-10, 1e-11 as the density RMS change falls
    below 1e-5, 1e-6, 1e-7 (never loosen); primLimit follows."""
    if prms <= 1e-5 and cutoff > 10.0 ** -8.5:
        cutoff = 1e-9
        prim_limit = min(cutoff, prim_limit)
    if prms <= 1e-6 and cutoff > 10.0 ** -9.5:
        cutoff = 1e-10
        prim_limit = min(cutoff, prim_limit)
    if prms <= 1e-7 and cutoff > 10.0 ** -10.5:
        cutoff = 1e-11
        prim_limit = min(cutoff, prim_limit)
    return cutoff, prim_limit


def level_shift(Fo, Cp_old, eps_old, nocc, it, errmax):
    """QUICK's level shift (quick_scf_module.f90:626-664): from DIIS cycle
    LSHIFT_CYCLE on, when the DIIS error exceeds LSHIFT_ERR and the previous
    HOMO-LUMO gap is below LSHIFT_GAP, the orthogonalised Fock matrix is
    taken into the previous orbital basis and the virtual diagonal raised by
    LSHIFT_GAP - gap.  Returns (matrix to diagonalise, back-rotation) or None."""
    if Cp_old is None or eps_old is None or nocc >= len(eps_old):
        return None
    gap = float(eps_old[nocc] - eps_old[nocc - 1])
    if not (it + 1 >= LSHIFT_CYCLE and errmax > LSHIFT_ERR and LSHIFT_GAP > gap):
        return None
    M = Cp_old.T @ Fo @ Cp_old
    idx = np.arange(nocc, M.shape[0])
    M[idx, idx] += LSHIFT_GAP - gap
    return M, Cp_old


def _orth(S):
    s, U = np.linalg.eigh(S)
    return (U / np.sqrt(s)) @ U.T


def _tighten(engine, prms):
    """Applies adjust_cutoff to the engine; True when the cutoff changed
    (the incremental Fock matrix then restarts from a full build)."""
    c0, p0 = engine.cutoff, engine.prim_limit
    c1, p1 = adjust_cutoff(prms, c0, p0)
    if (c1, p1) == (c0, p0):
        return False
    engine.set_cutoffs(c1, p1, engine.hyb)
    return True


class RHFResult(dict):
    __getattr__ = dict.__getitem__


def rhf(engine, nelec=None, maxiter=100, etol=1e-8, dtol=1e-6, ndiis=10, incremental=False, ncyc=3,
        guess=None, one_electron=None, verbose=False, cutoff_schedule=True):
    """Host-side RHF with GPU Fock builds.  ``cutoff_schedule``: QUICK's
    adjust_Cutoff tightening of the engine's integral cutoff as the density
    converges (a changed cutoff makes the next build a full one)."""
    mol = engine.basis.mol
    if nelec is None:
        nelec = mol.nelec()
    if nelec % 2:
        raise ValueError("RHF needs an even number of electrons")
    nocc = nelec // 2
    t0 = time.perf_counter()
    S, T, V = one_electron if one_electron is not None else engine.one_electron()
    H = T + V
    X = _orth(S)
    enuc = mol.nuclear_repulsion()

    Cp_old = eps_old = None

    def density(F, it=None, errmax=0.0):
        nonlocal Cp_old, eps_old
        Fo = X.T @ F @ X
        ls = level_shift(Fo, Cp_old, eps_old, nocc, it, errmax) if it is not None else None
        if ls is None:
            e, Cp = np.linalg.eigh(Fo)
        else:
            e, V = np.linalg.eigh(ls[0])
            Cp = ls[1] @ V
        Cp_old, eps_old = Cp, e
        C = X @ Cp
        return 2.0 * C[:, :nocc] @ C[:, :nocc].T, e

    P = guess if guess is not None else density(H)[0]
    fs, es, hist, ftimes = [], [], [], []
    E_old = 0.0
    F = None
    P_prev = None
    force_full = False
    for it in range(maxiter):
        tf = time.perf_counter()
        if incremental and it >= ncyc and F is not None and not force_full:
            F = F + engine.fock2e(P - P_prev)       # F_prev already holds H + G(P_prev)
        else:
            F = H + engine.fock2e(P)
        force_full = False
        ftimes.append(time.perf_counter() - tf)
        E = 0.5 * float(np.sum(P * (H + F))) + enuc
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        fs.append(F)
        es.append(err)
        fs, es = fs[-ndiis:], es[-ndiis:]
        k = len(fs)
        B = np.array([[float(np.sum(es[i] * es[j])) for j in range(k)] for i in range(k)])
        first, c = diis_coefficients(B)
        Fd = F if c is None else sum(ci * fi for ci, fi in zip(c, fs[first:]))
        Pn, eps = density(Fd, it, float(np.max(np.abs(err))))
        dP = float(np.sqrt(np.mean((Pn - P) ** 2)))
        hist.append(E)
        if verbose:
            print(f"{it:3d} {E:.12f} {E - E_old:+.3e} rmsdP {dP:.3e} fock {ftimes[-1] * 1e3:.1f} ms", flush=True)
        converged = abs(E - E_old) < etol and dP < dtol
        E_old = E
        P_prev = P
        P = Pn
        if converged:
            return RHFResult(energy=E, density=P_prev, iterations=it + 1, history=hist, fock_times=ftimes,
                             wall=time.perf_counter() - t0, converged=True, e_nuc=enuc)
        if cutoff_schedule:
            force_full = _tighten(engine, dP)
    return RHFResult(energy=E, density=P_prev, iterations=maxiter, history=hist, fock_times=ftimes,
                     wall=time.perf_counter() - t0, converged=False, e_nuc=enuc)


def uhf(engine, mult=None, nelec=None, maxiter=200, etol=1e-8, dtol=1e-6, ndiis=10, incremental=False, ncyc=3,
        one_electron=None, verbose=False):
    """Unrestricted HF with the open-shell GPU build (``FockEngine.fock2e_os``).
    n_alpha - n_beta = mult - 1; core-Hamiltonian guess.  With ``incremental``
    the builds after ``ncyc`` full ones add G(dPa, dPb) to the previous Fock
    matrices (deltaO, quick_uscf_operator_module.f90:71-86)."""
    mol = engine.basis.mol
    if nelec is None:
        nelec = mol.nelec()
    if mult is None:
        mult = 1 + nelec % 2
    if (nelec + mult - 1) % 2 or mult < 1 or mult - 1 > nelec:
        raise ValueError("multiplicity does not fit the electron count")
    na = (nelec + mult - 1) // 2
    nb = nelec - na
    t0 = time.perf_counter()
    S, T, V = one_electron if one_electron is not None else engine.one_electron()
    H = T + V
    X = _orth(S)
    enuc = mol.nuclear_repulsion()

    def density(F, nocc):
        _, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp[:, :nocc]
        return C @ C.T

    Pa, Pb = density(H, na), density(H, nb)
    fs, es, hist, ftimes = [], [], [], []
    E_old = 0.0
    Fa = Fb = None
    prev = None
    for it in range(maxiter):
        tf = time.perf_counter()
        if incremental and it >= ncyc and Fa is not None:
            Ga, Gb = engine.fock2e_os(Pa - prev[0], Pb - prev[1])
            Fa, Fb = Fa + Ga, Fb + Gb
        else:
            Ga, Gb = engine.fock2e_os(Pa, Pb)
            Fa, Fb = H + Ga, H + Gb
        ftimes.append(time.perf_counter() - tf)
        E = 0.5 * float(np.sum(Pa * (H + Fa)) + np.sum(Pb * (H + Fb))) + enuc
        err = np.concatenate([(X.T @ (F @ P @ S - S @ P @ F) @ X).ravel() for F, P in ((Fa, Pa), (Fb, Pb))])
        fs.append((Fa, Fb))
        es.append(err)
        fs, es = fs[-ndiis:], es[-ndiis:]
        k = len(fs)
        B = np.array([[float(np.dot(es[i], es[j])) for j in range(k)] for i in range(k)])
        first, c = diis_coefficients(B)
        if c is None:
            Fda, Fdb = Fa, Fb
        else:
            Fda = sum(ci * f[0] for ci, f in zip(c, fs[first:]))
            Fdb = sum(ci * f[1] for ci, f in zip(c, fs[first:]))
        Pna, Pnb = density(Fda, na), density(Fdb, nb)
        dP = max(float(np.sqrt(np.mean((Pna - Pa) ** 2))), float(np.sqrt(np.mean((Pnb - Pb) ** 2))))
        hist.append(E)
        if verbose:
            print(f"{it:3d} {E:.12f} {E - E_old:+.3e} rmsdP {dP:.3e} fock {ftimes[-1] * 1e3:.1f} ms", flush=True)
        converged = abs(E - E_old) < etol and dP < dtol
        E_old = E
        prev = (Pa, Pb)
        Pa, Pb = Pna, Pnb
        if converged:
            return RHFResult(energy=E, density=prev, iterations=it + 1, history=hist, fock_times=ftimes,
                             wall=time.perf_counter() - t0, converged=True, e_nuc=enuc, nalpha=na, nbeta=nb)
    return RHFResult(energy=E, density=prev, iterations=maxiter, history=hist, fock_times=ftimes,
                     wall=time.perf_counter() - t0, converged=False, e_nuc=enuc, nalpha=na, nbeta=nb)


def seeded_density(S, nocc, seed=20261202):
    """P = 2 C C^T with C (n x nocc) ~ N(0,1) from default_rng(seed), then
    S-orthonormalised, C <- C (C^T S C)^(-1/2): rank nocc, tr(PS) = 2 nocc."""
    n = S.shape[0]
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, nocc))
    w, U = np.linalg.eigh(C.T @ S @ C)
    C = C @ (U @ np.diag(w ** -0.5) @ U.T)
    return 2.0 * C @ C.T


def rhf_device(engine, nelec=None, maxiter=100, etol=1e-8, dtol=1e-6, ndiis=10, incremental=True, ncyc=3,
               rebuild=8, one_electron=None, device=None, verbose=False, cutoff_schedule=True):
    """RHF with every matrix on the GPU (torch FP64 tensors): Fock builds add
    G into the device F (``qf_fock_build_device``), DIIS and the eigensolver
    run on the device, and one small host transfer per iteration carries the
    energy, the DIIS overlaps and the density change.

    With ``incremental`` the build after ``ncyc`` full ones is QUICK's
    deltaO scheme (quick_scf_operator_module.f90:69-79): F = F_prev + G(P -
    P_prev), where the density screen sees |P - P_prev| and drops most
    quartets late in the SCF.  A full build is redone every ``rebuild``
    iterations and once more at convergence, and the SCF continues if that
    full-build energy moved by more than ``etol``.  DIIS uses QUICK's LSOLVE
    with the oldest-vector drop, level shifting and, with ``cutoff_schedule``,
    adjust_Cutoff (a changed cutoff forces a full build)."""
    import torch
    dev = device or torch.device("cuda:0")
    mol = engine.basis.mol
    if nelec is None:
        nelec = mol.nelec()
    if nelec % 2:
        raise ValueError("RHF needs an even number of electrons")
    nocc = nelec // 2
    t0 = time.perf_counter()
    S, T, V = one_electron if one_electron is not None else engine.one_electron()
    dd = dict(dtype=torch.float64, device=dev)
    S = torch.tensor(S, **dd)
    H = torch.tensor(T + V, **dd)
    s, U = torch.linalg.eigh(S)
    X = (U / torch.sqrt(s)) @ U.T
    enuc = mol.nuclear_repulsion()
    stream = torch.cuda.current_stream(dev).cuda_stream

    Cp_old = eps_old = None

    def density(F, it=None, errmax=0.0):
        nonlocal Cp_old, eps_old
        Fo = X.T @ F @ X
        shifted = False
        if it is not None and Cp_old is not None and nocc < len(eps_old):
            gap = float(eps_old[nocc] - eps_old[nocc - 1])
            if it + 1 >= LSHIFT_CYCLE and errmax > LSHIFT_ERR and LSHIFT_GAP > gap:
                M = Cp_old.T @ Fo @ Cp_old
                d = torch.arange(nocc, M.shape[0], device=M.device)
                M[d, d] += LSHIFT_GAP - gap
                e, V = torch.linalg.eigh(M)
                Cp = Cp_old @ V
                shifted = True
        if not shifted:
            e, Cp = torch.linalg.eigh(Fo)
        Cp_old, eps_old = Cp, e.cpu().numpy()
        C = X @ Cp[:, :nocc]
        return 2.0 * C @ C.T

    P = density(H)
    F = torch.empty_like(H)
    dP = torch.empty_like(H)
    P_prev = None
    fs, es = [], []
    B = np.zeros((0, 0))
    E_old = 0.0
    hist, ftimes, full_flags = [], [], []
    it_full = 0
    converged = False
    force_full = False
    E = 0.0
    for it in range(maxiter):
        tf = time.perf_counter()
        full = (not incremental) or it < ncyc or P_prev is None or (it - it_full) >= rebuild or force_full
        force_full = False
        if full:
            F.copy_(H)
            engine.fock2e_device(P.data_ptr(), F.data_ptr(), stream)
            it_full = it
        else:
            torch.sub(P, P_prev, out=dP)
            engine.fock2e_device(dP.data_ptr(), F.data_ptr(), stream)   # F already holds H + G(P_prev)
        full_flags.append(full)
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        fs.append(F.clone())
        es.append(err)
        if len(fs) > ndiis:
            fs.pop(0)
            es.pop(0)
            B = B[1:, 1:]
        Ed = 0.5 * torch.sum(P * (H + F))
        row = torch.stack([Ed, torch.max(torch.abs(err))] + [torch.sum(err * e) for e in es]).cpu().numpy()
        errmax = float(row[1])
        row = np.concatenate([row[:1], row[2:]])
        ftimes.append(time.perf_counter() - tf)            # build + DIIS row, to the one host sync
        E = float(row[0]) + enuc
        k = len(es)
        Bn = np.zeros((k, k))
        Bn[:k - 1, :k - 1] = B
        Bn[k - 1, :] = row[1:]
        Bn[:, k - 1] = row[1:]
        B = Bn
        first, c = diis_coefficients(B)
        if c is None:
            Fd = F
        else:
            Fd = fs[first] * float(c[0])
            for ci, fi in zip(c[1:], fs[first + 1:]):
                Fd = Fd + float(ci) * fi
        Pn = density(Fd, it, errmax)
        rms = float(torch.sqrt(torch.mean((Pn - P) ** 2)))
        hist.append(E)
        if verbose:
            print(f"{it:3d} {E:.12f} {E - E_old:+.3e} rmsdP {rms:.3e} {'full' if full else 'delta'} "
                  f"{ftimes[-1] * 1e3:.1f} ms", flush=True)
        done = abs(E - E_old) < etol and rms < dtol
        E_old = E
        if done:
            if full:
                converged = True
                break
            it_full = it - rebuild          # force a full build next: confirm on the full Fock matrix
        elif cutoff_schedule and _tighten(engine, rms):
            force_full = True
        P_prev = P
        P = Pn
    torch.cuda.synchronize(dev)
    return RHFResult(energy=E, density=P.cpu().numpy(), iterations=len(hist), history=hist, fock_times=ftimes,
                     full_builds=int(sum(full_flags)), wall=time.perf_counter() - t0, converged=converged,
                     e_nuc=enuc)
