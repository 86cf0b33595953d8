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


def _orth(S):
    s, U = np.linalg.eigh(S)
    return (U / np.sqrt(s)) @ U.T


class RHFResult(dict):
    __getattr__ = dict.__getitem__


def rhf(engine, nelec=None, maxiter=100, etol=1e-8, dtol=1e-6, ndiis=10, incremental=False, ncyc=3,
        guess=None, one_electron=None, verbose=False):
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

    def density(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        return 2.0 * C[:, :nocc] @ C[:, :nocc].T, e

    P = guess if guess is not None else density(H)[0]
    fs, es, hist, ftimes = [], [], [], []
    E_old = 0.0
    F = None
    P_prev = None
    for it in range(maxiter):
        tf = time.perf_counter()
        if incremental and it >= ncyc and F is not None:
            F = F + engine.fock2e(P - P_prev)       # F_prev already holds H + G(P_prev)
        else:
            F = H + engine.fock2e(P)
        ftimes.append(time.perf_counter() - tf)
        E = 0.5 * float(np.sum(P * (H + F))) + enuc
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        fs.append(F)
        es.append(err)
        fs, es = fs[-ndiis:], es[-ndiis:]
        k = len(fs)
        B = -np.ones((k + 1, k + 1))
        B[k, k] = 0.0
        for i in range(k):
            for j in range(k):
                B[i, j] = float(np.sum(es[i] * es[j]))
        rhs = np.zeros(k + 1)
        rhs[k] = -1.0
        try:
            c = np.linalg.solve(B, rhs)[:k]
            Fd = sum(ci * fi for ci, fi in zip(c, fs))
        except np.linalg.LinAlgError:
            Fd = F
        Pn, eps = density(Fd)
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
        B = -np.ones((k + 1, k + 1))
        B[k, k] = 0.0
        for i in range(k):
            for j in range(k):
                B[i, j] = float(np.dot(es[i], es[j]))
        rhs = np.zeros(k + 1)
        rhs[k] = -1.0
        try:
            c = np.linalg.solve(B, rhs)[:k]
            Fda = sum(ci * f[0] for ci, f in zip(c, fs))
            Fdb = sum(ci * f[1] for ci, f in zip(c, fs))
        except np.linalg.LinAlgError:
            Fda, Fdb = Fa, Fb
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
               rebuild=8, one_electron=None, device=None, verbose=False):
    """RHF with every matrix on the GPU (torch FP64 tensors): Fock builds add
    G into the device F (``qf_fock_build_device``), DIIS and the eigensolver
    run on the device, and one small host transfer per iteration carries the
    energy, the DIIS overlaps and the density change.

    With ``incremental`` the build after ``ncyc`` full ones is QUICK's
    deltaO scheme (quick_scf_operator_module.f90:69-79): F = F_prev + G(P -
    P_prev), where the density screen sees |P - P_prev| and drops most
    quartets late in the SCF.  A full build is redone every ``rebuild``
    iterations and once more at convergence, and the SCF continues if that
    full-build energy moved by more than ``etol``."""
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

    def density(F):
        _, Cp = torch.linalg.eigh(X.T @ F @ X)
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
    E = 0.0
    for it in range(maxiter):
        tf = time.perf_counter()
        full = (not incremental) or it < ncyc or P_prev is None or (it - it_full) >= rebuild
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
        row = torch.stack([Ed] + [torch.sum(err * e) for e in es]).cpu().numpy()
        ftimes.append(time.perf_counter() - tf)            # build + DIIS row, to the one host sync
        E = float(row[0]) + enuc
        k = len(es)
        Bn = np.zeros((k, k))
        Bn[:k - 1, :k - 1] = B
        Bn[k - 1, :] = row[1:]
        Bn[:, k - 1] = row[1:]
        B = Bn
        A = -np.ones((k + 1, k + 1))
        A[:k, :k] = B
        A[k, k] = 0.0
        rhs = np.zeros(k + 1)
        rhs[k] = -1.0
        try:
            c = np.linalg.solve(A, rhs)[:k]
            Fd = fs[0] * float(c[0])
            for ci, fi in zip(c[1:], fs[1:]):
                Fd = Fd + float(ci) * fi
        except np.linalg.LinAlgError:
            Fd = F
        Pn = density(Fd)
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
        P_prev = P
        P = Pn
    torch.cuda.synchronize(dev)
    return RHFResult(energy=E, density=P.cpu().numpy(), iterations=len(hist), history=hist, fock_times=ftimes,
                     full_builds=int(sum(full_flags)), wall=time.perf_counter() - t0, converged=converged,
                     e_nuc=enuc)
