"""Continuous-batching SCF server on one GPU.

Requests (a molecule and a basis) arrive over time.  Each request becomes a
device-resident RHF job: S, H = T + V, X = S^-1/2, P, F and the DIIS history
live on the GPU as FP64 tensors for the job's whole life.  A server step is
one SCF iteration of every active job:

1. per job the input density of the build: P after F = H (a full build), or
   P - P_prev with F keeping H + G(P_prev) (QUICK's deltaO,
   quick_scf_operator_module.f90:69-79; a full build every ``rebuild``
   iterations, after a cutoff change and to confirm convergence); then ONE
   batched Fock build adds G into every F (one launch per quartet class for
   all jobs, replayed from a HIP graph, ``qf_batch_create_ext`` over the
   jobs' own tensors);
2. per job the DIIS error X^T (F P S - S P F) X and energy, one host transfer
   for the small DIIS systems of all jobs (QUICK's LSOLVE with the
   oldest-vector drop), the extrapolated Fock matrix and diagonalisation
   (jobs of equal size share one batched ``eigh``);
3. QUICK's adjust_Cutoff tightens a job's integral cutoff as its density
   converges (quick_method_module.f90:1105-1128);
4. a converged job leaves: its input density is set to zero, so its quartets
   all fail the density screen, and it stays in the batch (and its graph)
   until the next rebuild, when its engine and tensors are released.

New requests are prepared (engine, Schwarz bounds, pair and task tables,
one-electron integrals) on a worker thread while the steps run, and join at
the next rebuild of the batch: at once when no job is active, otherwise
after ``regroup`` steps.  Admission is by measured device bytes (the
engine's own tables plus the job's tensors) against a budget taken from the
device's free memory.  A job's result does not depend on which other jobs
share its batch: the batched build gives every molecule exactly its own G,
and everything else reads only the job's own tensors.  ``FockServer.run``
reports per-request latency (arrival to convergence) and throughput.
"""
from __future__ import annotations

import heapq
import queue
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import basis as B
from .engine import FockBatch, FockEngine
from .scf import adjust_cutoff, diis_coefficients


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
    full_builds: int = 0
    cutoff: float = 0.0

    @property
    def latency(self):
        return self.finished - self.arrival


class ScfJob:
    """Device-resident RHF state of one request."""

    def __init__(self, req: Request, device, precision="mixed", cutoff=1e-8, ndiis=10, etol=1e-8, dtol=1e-6,
                 incremental=True, ncyc=3, rebuild=8, cutoff_schedule=True):
        import torch
        self.torch = torch
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
        self.Pin = torch.zeros((n, n), **dd)        # input density of the batched build
        self.P_prev = None
        self.enuc = req.mol.nuclear_repulsion()
        self.ndiis, self.etol, self.dtol = ndiis, etol, dtol
        self.incremental, self.ncyc, self.rebuild = incremental, ncyc, rebuild
        self.cutoff_schedule = cutoff_schedule
        self.fs, self.es = [], []
        self.B = np.zeros((0, 0))
        self.E_old = 0.0
        self.energy = 0.0
        self.it = 0
        self.it_full = 0
        self.full = True
        self.force_full = False
        self.full_builds = 0
        self.want_cutoff = None
        self.done = False
        self.converged = False
        self.P.copy_(self._density(self.H))          # core guess
        torch.cuda.synchronize(device)
        # measured device bytes: the engine's tables and buffers plus this job's tensors
        self.bytes = int(self.engine.stats()["device_bytes"]) + 6 * n * n * 8

    def _density(self, F):
        Fp = self.X.T @ F @ self.X
        _, Cp = self.torch.linalg.eigh(Fp)
        C = self.X @ Cp[:, :self.nocc]
        return 2.0 * C @ C.T

    # --- step, part 0 (device): input of the batched build -----------------
    def begin_step(self):
        full = (not self.incremental or self.it < self.ncyc or self.P_prev is None or
                self.it - self.it_full >= self.rebuild or self.force_full)
        self.force_full = False
        if full:
            self.F.copy_(self.H)
            self.Pin.copy_(self.P)
            self.it_full = self.it
            self.full_builds += 1
        else:
            self.torch.sub(self.P, self.P_prev, out=self.Pin)   # F already holds H + G(P_prev)
        self.full = full

    def retire(self):
        """Leave the batch without rebuilding it: a zero density fails every
        quartet's density screen."""
        self.Pin.zero_()
        self.done = True

    # --- step, part 1 (device): error vector, energy, DIIS overlaps -------
    def diis_row(self):
        torch = self.torch
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

    # --- step, part 2 (host): DIIS coefficients (LSOLVE, oldest dropped) ---
    def diis_extrapolated(self, row):
        E = float(row[0]) + self.enuc
        dots = row[1:]
        k = len(dots)
        Bn = np.zeros((k, k))
        Bn[:k - 1, :k - 1] = self.B
        Bn[k - 1, :] = dots
        Bn[:, k - 1] = dots
        self.B = Bn
        self.energy = E
        first, c = diis_coefficients(Bn)
        if c is None:
            return self.F
        Fd = self.fs[first] * float(c[0])
        for ci, fi in zip(c[1:], self.fs[first + 1:]):
            Fd = Fd + float(ci) * fi
        return Fd

    # --- step, part 3 (host): convergence and the cutoff schedule ---------
    def finish_step(self, Pnew, rms):
        self.it += 1
        done = abs(self.energy - self.E_old) < self.etol and rms < self.dtol
        self.E_old = self.energy
        if done and self.full:
            self.converged = True
            return True
        if done:
            self.force_full = True          # confirm on a full build
        elif self.cutoff_schedule:
            c1, p1 = adjust_cutoff(rms, self.engine.cutoff, self.engine.prim_limit)
            if (c1, p1) != (self.engine.cutoff, self.engine.prim_limit):
                self.want_cutoff = (c1, p1)
        self.P_prev = self.P.clone()
        self.P.copy_(Pnew)
        return False

    def apply_cutoff(self):
        """A changed cutoff re-prepares the engine (the batch is rebuilt
        around it) and restarts the incremental Fock matrix from a full build."""
        if self.want_cutoff is None:
            return False
        c1, p1 = self.want_cutoff
        self.want_cutoff = None
        self.engine.set_cutoffs(c1, p1, self.engine.hyb)
        self.engine.prepare()
        self.force_full = True
        return True

    def close(self):
        self.engine.close()
        self.fs, self.es = [], []


def footprint(req, ndiis=10, replicas=8):
    """Device bytes of a job before it is prepared (admission estimate):
    G replicas, P/F/S/H/X and work matrices, the DIIS history (2 x ndiis
    matrices) and the pair tables (a few nbasis^2 doubles for these bases)."""
    n = B.build_basis(req.mol, req.basis).nbasis
    return (replicas + 12 + 2 * ndiis) * n * n * 8


class FockServer:
    """Runs SCF requests with continuous batching: a worker thread prepares
    arriving requests, the step loop admits prepared ones by measured bytes
    and slot count, and all active jobs iterate together."""

    def __init__(self, device=None, hbm_budget=None, max_slots=64, precision="mixed", cutoff=1e-8,
                 replicas=2, etol=1e-8, dtol=1e-6, maxiter=60, regroup=4, graph=True, incremental=True,
                 cutoff_schedule=True):
        import torch
        self.torch = torch
        self.device = device or torch.device("cuda:0")
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = 0.8 * free
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.precision = precision
        self.cutoff = cutoff
        self.replicas = replicas
        self.etol, self.dtol, self.maxiter = etol, dtol, maxiter
        self.regroup = int(regroup)
        self.graph = graph
        self.incremental = incremental
        self.cutoff_schedule = cutoff_schedule
        self.stream = torch.cuda.current_stream(self.device)

    def _make_job(self, req):
        return ScfJob(req, self.device, self.precision, self.cutoff, etol=self.etol, dtol=self.dtol,
                      incremental=self.incremental, cutoff_schedule=self.cutoff_schedule)

    def run(self, requests, log=None):
        torch = self.torch
        arrivals = sorted(requests, key=lambda r: r.arrival)
        todo = queue.Queue()
        ready = []                  # prepared jobs: (arrival, rid, job)
        ready_lock = threading.Lock()
        # HIP does not keep another thread's work out of a stream capture:
        # the worker's preparation and the build that captures a new batch's
        # graph take turns
        capture_lock = threading.Lock()
        errors = []

        def worker():
            torch.cuda.set_device(self.device)
            while True:
                req = todo.get()
                if req is None:
                    return
                try:
                    with capture_lock:
                        job = self._make_job(req)
                except Exception as exc:          # surfaced by the step loop
                    errors.append(exc)
                    return
                with ready_lock:
                    heapq.heappush(ready, (req.arrival, req.rid, job))

        th = threading.Thread(target=worker, daemon=True)
        th.start()
        active = []                 # jobs in the batch (retired ones stay until the next rebuild)
        results = []
        batch = None
        used = 0
        t0 = time.perf_counter()
        qi = 0
        steps = since = rebuilds = 0
        try:
            while qi < len(arrivals) or len(results) < len(arrivals):
                if errors:
                    raise errors[0]
                now = time.perf_counter() - t0
                while qi < len(arrivals) and arrivals[qi].arrival <= now:
                    todo.put(arrivals[qi])
                    qi += 1
                live = [j for j in active if not j.done]
                with ready_lock:
                    have_ready = bool(ready)
                wants = any(j.want_cutoff is not None for j in live)
                regroup = batch is None or wants or (have_ready and (not live or since >= self.regroup)) or \
                    (len(active) > 2 * max(1, len(live)))
                if regroup:
                    # release retired jobs, apply pending cutoff changes, admit prepared jobs
                    if batch is not None:
                        torch.cuda.synchronize(self.device)
                        batch.close()
                        batch = None
                    for j in active:
                        if j.done:
                            used -= j.bytes
                            j.close()
                    active = live
                    for j in active:
                        j.apply_cutoff()
                    with ready_lock:
                        while ready and len(active) < self.max_slots:
                            job = ready[0][2]
                            if active and used + job.bytes > self.hbm_budget:
                                break
                            heapq.heappop(ready)
                            job.admitted = time.perf_counter() - t0
                            used += job.bytes
                            active.append(job)
                    if active:
                        batch = FockBatch([j.engine for j in active], replicas=self.replicas, graph=self.graph,
                                          density_ptrs=[j.Pin.data_ptr() for j in active],
                                          fock_ptrs=[j.F.data_ptr() for j in active])
                        rebuilds += 1
                    since = 0
                live = [j for j in active if not j.done]
                if not live:
                    if len(results) < len(arrivals):
                        time.sleep(0.0005)
                    continue
                # 1. the build input of every live job, then one batched build
                for j in live:
                    j.begin_step()
                if since == 0 and self.graph:
                    with capture_lock:              # the first build of a batch captures its graph
                        batch.build(self.stream.cuda_stream)
                else:
                    batch.build(self.stream.cuda_stream)
                # 2. DIIS rows on the device, one transfer for all jobs
                rows = [j.diis_row() for j in live]
                flat = torch.cat(rows).cpu().numpy()
                off = 0
                Fds = []
                for j, r in zip(live, rows):
                    Fds.append(j.diis_extrapolated(flat[off:off + r.numel()]))
                    off += r.numel()
                # diagonalise, jobs of equal size together
                Pnew = [None] * len(live)
                by_n = {}
                for k, j in enumerate(live):
                    by_n.setdefault(j.n, []).append(k)
                for n, ks in by_n.items():
                    X = torch.stack([live[k].X for k in ks])
                    Fp = X.transpose(1, 2) @ torch.stack([Fds[k] for k in ks]) @ X
                    _, Cp = torch.linalg.eigh(Fp)
                    for q, k in enumerate(ks):
                        C = live[k].X @ Cp[q][:, :live[k].nocc]
                        Pnew[k] = 2.0 * C @ C.T
                rms = torch.stack([torch.sqrt(torch.mean((Pn - j.P) ** 2)) for Pn, j in zip(Pnew, live)])
                rms = rms.cpu().numpy()
                for k, j in enumerate(live):
                    conv = j.finish_step(Pnew[k], float(rms[k]))
                    if conv or j.it >= self.maxiter:
                        j.converged = bool(conv)
                        fin = time.perf_counter() - t0
                        results.append(Result(j.req.rid, j.energy, j.it, j.converged, j.req.arrival, j.admitted,
                                              fin, j.n, j.full_builds, j.engine.cutoff))
                        j.retire()
                steps += 1
                since += 1
                if log:
                    log(f"step {steps}: active {len(live)} ready {len(ready)} done {len(results)}")
        finally:
            todo.put(None)
            th.join()
            torch.cuda.synchronize(self.device)
            if batch is not None:
                batch.close()
            for j in active:
                j.close()
            for _, _, j in ready:
                j.close()
        wall = time.perf_counter() - t0
        return results, {"wall_s": wall, "steps": steps, "rebuilds": rebuilds}


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
