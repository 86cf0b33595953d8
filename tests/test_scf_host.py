"""QUICK's SCF safety nets restated in quick_amd.scf: LSOLVE (Lsolve.f90),
the DIIS oldest-vector drop (quick_scf_module.f90:540-586), adjust_Cutoff
(quick_method_module.f90:1105-1128) and level shifting (:626-664)."""
import numpy as np
import pytest

from quick_amd import scf


def test_lsolve_matches_numpy_and_flags_singular_systems():
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 11):
        A = rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n)
        assert np.allclose(scf.lsolve(A, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    A = rng.standard_normal((4, 4))
    A[3] = A[0] + A[1]                     # rank 3
    assert scf.lsolve(A, np.ones(4)) is None
    assert scf.lsolve(np.zeros((1, 1)), np.ones(1)) is None
    # a pivot below thresh * mean|A| fails, one above passes
    A = np.diag([1.0, 1e-13])
    assert scf.lsolve(A, np.ones(2), thresh=1e-10) is None
    assert scf.lsolve(np.diag([1.0, 1e-9]), np.ones(2), thresh=1e-10) is not None


def test_diis_drops_oldest_vectors_until_solvable():
    rng = np.random.default_rng(3)
    e = [rng.standard_normal(6) for _ in range(3)]
    e = [e[0], e[0], e[1], e[2]]           # the two oldest vectors are equal: singular B
    B = np.array([[x @ y for y in e] for x in e])
    first, c = scf.diis_coefficients(B)
    assert first == 1 and c is not None and abs(c.sum() - 1.0) < 1e-12
    # the coefficients minimise |sum c e| over the kept vectors
    Bk = B[1:, 1:]
    A = -np.ones((4, 4))
    A[:3, :3] = Bk
    A[3, 3] = 0.0
    assert np.allclose(c, np.linalg.solve(A, np.r_[0, 0, 0, -1.0])[:3])
    # all-zero errors: only the newest vector survives (the plain Fock matrix)
    first, c = scf.diis_coefficients(np.zeros((3, 3)))
    assert first == 2 and np.allclose(c, [1.0])


def test_adjust_cutoff_schedule():
    assert scf.adjust_cutoff(1e-4, 1e-7, 1e-8) == (1e-7, 1e-8)
    assert scf.adjust_cutoff(1e-5, 1e-7, 1e-8) == (1e-9, 1e-9)
    assert scf.adjust_cutoff(5e-7, 1e-7, 1e-8) == (1e-10, 1e-10)
    assert scf.adjust_cutoff(1e-8, 1e-7, 1e-8) == (1e-11, 1e-11)
    assert scf.adjust_cutoff(1e-8, 1e-12, 1e-13) == (1e-12, 1e-13)   # never loosens
    assert scf.adjust_cutoff(1e-8, 1e-11, 1e-12) == (1e-11, 1e-12)


def test_level_shift_raises_virtuals_in_the_old_orbital_basis():
    rng = np.random.default_rng(5)
    n, nocc = 6, 2
    M = rng.standard_normal((n, n))
    Fo = M + M.T
    eps, C = np.linalg.eigh(Fo)
    eps_small_gap = eps.copy()
    eps_small_gap[nocc] = eps_small_gap[nocc - 1] + 0.05
    assert scf.level_shift(Fo, C, eps_small_gap, nocc, 0, 1.0) is None     # before cycle 3
    assert scf.level_shift(Fo, C, eps_small_gap, nocc, 5, 0.01) is None    # error too small
    assert scf.level_shift(Fo, C, eps, nocc, 5, 1.0) is None if eps[nocc] - eps[nocc - 1] >= 0.2 else True
    Ms, R = scf.level_shift(Fo, C, eps_small_gap, nocc, 5, 1.0)
    shift = 0.2 - 0.05
    # in the old orbital basis (the eigenbasis of Fo) the shifted matrix is
    # diag(eps) with the virtual block raised by the shift
    assert np.allclose(Ms, np.diag(np.r_[eps[:nocc], eps[nocc:] + shift]), atol=1e-12)
    assert R is C
