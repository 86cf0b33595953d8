"""Continuous-batching SCF server on one GPU.

Requests (a molecule and a basis) arrive over time.  Each admitted request
becomes a device-resident RHF job: S, H = T + V, X = S^-1/2, P, F and the DIIS
history live on the GPU as FP64 tensors for the job's whole life.  A step of
the server is one SCF iteration of every active job:

1. F_i = H_i, then ONE batched Fock build adds G(P_i) into every F_i (one
   launch per quartet class for all jobs, ``qf_batch_create_ext`` over the
   jobs' own P/F tensors);
2. per job the DIIS error X^T (F P S - S P F) X and energy, one host transfer
   for the small DIIS systems of all jobs, the extrapolated Fock matrix, and
   diagonalisation (jobs of equal size share one batched ``eigh``);
3. converged jobs leave and free their memory; queued requests are admitted
   while the HBM budget and the slot count allow, so new work joins the
   running batch at the next step (continuous batching).

A job's result does not depend on which other jobs share its batch: the
batched build gives every molecule exactly its own G(P), and the SCF of each
job only reads its own tensors.  ``FockServer.run`` reports per-request
latency (arrival to convergence) and throughput.

The SCF restates QUICK's electdiis control flow (quick_scf_module.f90:344-760)
as in ``quick_amd.scf.rhf``: F = H + G(P), E = 1/2 sum P (H + F) + E_nuc,
Pulay DIIS with up to 10 vectors, P = 2 C_occ C_occ^T.
"""
from __future__ import annotations

import heapq
import time
from dataclasses import dataclass, field

import numpy as np

from . import basis as B
from .engine import FockBatch, FockEngine


@dataclass
class Request:
    rid: int
    mol: B.Molecule
    basis: str
    arrival: float = 0.0          # seconds after the start of the run


@dataclass
class Result:
    rid: int
    energy: float
    iterations: int
    converged: bool
    arrival: float
    admitted: float
    finished: float
    nbasis: int

    @property
    def latency(self):
        return self.finished - self.arrival


class ScfJob:
    """Device-resident RHF state of one request."""

    def __init__(self, req: Request, device, precision="mixed", cutoff=1e-10, ndiis=10, etol=1e-8, dtol=1e-6):
        import torch
        self.req = req
        self.bs = B.build_basis(req.mol, req.basis)
        self.n = n = self.bs.nbasis
        nelec = req.mol.nelec()
        if nelec % 2:
            raise ValueError("RHF needs an even number of electrons")
        self.nocc = nelec // 2
        self.engine = FockEngine(self.bs, device=device.index or 0, cutoff=cutoff, precision=precision)
        self.engine.prepare()
        S, T, V = self.engine.one_electron()
        dd = dict(dtype=torch.float64, device=device)
        self.S = torch.tensor(S, **dd)
        self.H = torch.tensor(T + V, **dd)
        s, U = torch.linalg.eigh(self.S)
        self.X = (U / torch.sqrt(s)) @ U.T
        self.F = self.H.clone()
        self.P = torch.zeros((n, n), **dd)
        self.enuc = req.mol.nuclear_repulsion()
        self.ndiis, self.etol, self.dtol = ndiis, etol, dtol
        self.fs, self.es = [], []
        self.B = np.zeros((0, 0))
        self.E_old = 0.0
        self.energy = 0.0
        self.it = 0
        self.done = False
        self.converged = False
        self.P.copy_(self._density(self.H))          # core guess
        self.bytes = 0

    def _density(self, F):
        Fp = self.X.T @ F @ self.X
        _, Cp = __import__("torch").linalg.eigh(Fp)
        C = self.X @ Cp[:, :self.nocc]
        return 2.0 * C @ C.T

    # --- step, part 1 (device): error vector, energy, DIIS overlaps -------
    def diis_row(self):
        import torch
        F, P, S, X = self.F, self.P, self.S, self.X
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        self.fs.append(F.clone())
        self.es.append(err)
        if len(self.fs) > self.ndiis:
            self.fs.pop(0)
            self.es.pop(0)
            self.B = self.B[1:, 1:]
        E = 0.5 * torch.sum(P * (self.H + F))
        dots = torch.stack([torch.sum(err * e) for e in self.es])
        return torch.cat([E.reshape(1), dots])

    # --- step, part 2 (host): DIIS coefficients ----------------------------
    def diis_coeffs(self, row):
        E = float(row[0]) + self.enuc
        dots = row[1:]
        k = len(dots)
        Bn = np.zeros((k, k))
        Bn[:k - 1, :k - 1] = self.B
        Bn[k - 1, :] = dots
        Bn[:, k - 1] = dots
        self.B = Bn
        A = -np.ones((k + 1, k + 1))
        A[:k, :k] = Bn
        A[k, k] = 0.0
        rhs = np.zeros(k + 1)
        rhs[k] = -1.0
        try:
            c = np.linalg.solve(A, rhs)[:k]
        except np.linalg.LinAlgError:
            c = np.zeros(k)
            c[-1] = 1.0
        self.energy = E
        return c

    def extrapolated(self, c):
        Fd = self.fs[0] * float(c[0])
        for ci, fi in zip(c[1:], self.fs[1:]):
            Fd = Fd + float(ci) * fi
        return Fd


def footprint(req, ndiis=10, replicas=8):
    """Device bytes a job holds: G replicas, P/F/S/H/X and work matrices,
    the DIIS history (2 x ndiis matrices) and the pair tables (bounded by
    a few nbasis^2 doubles for these basis sets)."""
    n = B.build_basis(req.mol, req.basis).nbasis
    return (replicas + 12 + 2 * ndiis) * n * n * 8


class FockServer:
    """Admits SCF requests by HBM budget and slot count and runs all active
    jobs' iterations together (continuous batching)."""

    def __init__(self, device=None, hbm_budget=8 << 30, max_slots=64, precision="mixed", cutoff=1e-10,
                 replicas=2, etol=1e-8, dtol=1e-6, maxiter=60):
        import torch
        self.torch = torch
        self.device = device or torch.device("cuda:0")
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.precision = precision
        self.cutoff = cutoff
        self.replicas = replicas
        self.etol, self.dtol, self.maxiter = etol, dtol, maxiter
        self.stream = torch.cuda.current_stream(self.device)

    def run(self, requests, log=None):
        torch = self.torch
        queue = sorted(requests, key=lambda r: r.arrival)
        pending = []            # arrived, not admitted: (arrival, rid, req)
        active = []
        results = []
        batch = None
        batch_key = None
        used = 0
        t0 = time.perf_counter()
        qi = 0
        steps = 0
        while qi < len(queue) or pending or active:
            now = time.perf_counter() - t0
            while qi < len(queue) and queue[qi].arrival <= now:
                heapq.heappush(pending, (queue[qi].arrival, queue[qi].rid, queue[qi]))
                qi += 1
            # admission: oldest first while memory and slots allow
            while pending and len(active) < self.max_slots:
                req = pending[0][2]
                if active and used + footprint(req) > self.hbm_budget:
                    break
                heapq.heappop(pending)
                job = ScfJob(req, self.device, self.precision, self.cutoff, etol=self.etol, dtol=self.dtol)
                job.bytes = footprint(req)
                job.admitted = time.perf_counter() - t0
                used += job.bytes
                active.append(job)
            if not active:
                # idle until the next arrival
                if qi < len(queue):
                    time.sleep(max(0.0, queue[qi].arrival - (time.perf_counter() - t0)))
                continue
            key = tuple(id(j) for j in active)
            if key != batch_key:
                if batch is not None:
                    batch.close()
                batch = FockBatch([j.engine for j in active], replicas=self.replicas,
                                  density_ptrs=[j.P.data_ptr() for j in active],
                                  fock_ptrs=[j.F.data_ptr() for j in active])
                batch_key = key
            # 1. F = H + G(P) for every job, one batched build
            for j in active:
                j.F.copy_(j.H)
            batch.build(self.stream.cuda_stream)
            # 2. DIIS rows on the device, one transfer for all jobs
            rows = [j.diis_row() for j in active]
            flat = torch.cat(rows).cpu().numpy()
            off = 0
            Fds = []
            for j, r in zip(active, rows):
                c = j.diis_coeffs(flat[off:off + r.numel()])
                off += r.numel()
                Fds.append(j.extrapolated(c))
            # diagonalise, jobs of equal size together
            Pnew = [None] * len(active)
            by_n = {}
            for k, j in enumerate(active):
                by_n.setdefault(j.n, []).append(k)
            for n, ks in by_n.items():
                X = torch.stack([active[k].X for k in ks])
                Fp = X.transpose(1, 2) @ torch.stack([Fds[k] for k in ks]) @ X
                _, Cp = torch.linalg.eigh(Fp)
                for q, k in enumerate(ks):
                    C = active[k].X @ Cp[q][:, :active[k].nocc]
                    Pnew[k] = 2.0 * C @ C.T
            dP = torch.stack([torch.sqrt(torch.mean((Pn - j.P) ** 2)) for Pn, j in zip(Pnew, active)]).cpu().numpy()
            still = []
            for k, j in enumerate(active):
                j.it += 1
                conv = abs(j.energy - j.E_old) < j.etol and dP[k] < j.dtol
                j.E_old = j.energy
                if conv or j.it >= self.maxiter:
                    j.converged = bool(conv)
                    fin = time.perf_counter() - t0
                    results.append(Result(j.req.rid, j.energy, j.it, j.converged, j.req.arrival, j.admitted, fin, j.n))
                    used -= j.bytes
                else:
                    j.P.copy_(Pnew[k])
                    still.append(j)
            if len(still) != len(active):
                torch.cuda.synchronize(self.device)
                batch.close()
                batch, batch_key = None, None
                for j in active:
                    if j not in still:
                        j.engine.close()
            active = still
            steps += 1
            if log:
                log(f"step {steps}: active {len(active)} pending {len(pending)} done {len(results)}")
        if batch is not None:
            batch.close()
        wall = time.perf_counter() - t0
        return results, {"wall_s": wall, "steps": steps}


def request_stream(count, rate, seed=20261202, bases=("6-31G", "6-31G*", "def2-SVP"), max_atoms=60):
    """Poisson arrivals at ``rate`` requests/s of molecules from the batch
    generator (water clusters 1-16, n-alkanes 1-12; 3-60 atoms), each with a
    basis drawn uniformly from ``bases``."""
    from . import molecules as M
    rng = np.random.default_rng(seed)
    pool = [m for m in M.mixed_set(28) if 3 <= m.natom <= max_atoms]
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        mol = pool[int(rng.integers(len(pool)))]
        out.append(Request(i, mol, bases[int(rng.integers(len(bases)))], t))
    return out


def summarize(results, wall):
    lat = np.array([r.latency for r in results])
    return {"requests": len(results), "requests_per_s": len(results) / wall,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "converged": int(sum(r.converged for r in results))}
