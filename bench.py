"""Benchmark: two-electron Fock builds per second on one MI355X.

Workload (FP64): RHF/6-31G(d) on a seeded 32-water cluster (96 atoms, 608
Cartesian functions), QUICK's screening at integral cutoff 1e-11 and
primitive limit 1e-12 (the values QUICK's adjust_Cutoff ends on near
convergence).  The density is the converged SCF density (SCF run first, its
wall time and energy reported beside the throughput), resident in HBM.  A
step is one full Fock build G(P) (screen, all class kernels, symmetrise,
F += G).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--precision fp64|mixed]

Prints one JSON line.  The CPU baseline is the oracle (the C restatement of
QUICK's CPU get2e) running full builds of the same workload with OpenMP on
the host cores (median of --cpu-builds), plus a bounded single-thread sample.

Other workloads (--workload): batch (256 uneven molecules per build), serve
(continuous-batching SCF server), gly10 (split precision), uhf (open-shell
build) and quick_tests (QUICK's own RHF regression inputs against the build
times in their saved outputs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fock builds/s and SCF wall time vs plain FP64 at ≤1e-6 Eh; serving p50/p99"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--nwater", type=int, default=32)
    ap.add_argument("--cutoff", type=float, default=1e-11)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--cpu-builds", type=int, default=5, help="oracle builds timed for the CPU baseline (median)")
    ap.add_argument("--workload", default="h2o32", choices=["h2o32", "batch", "serve", "gly10", "uhf", "quick_tests"],
                    help="h2o32: (H2O)32 6-31G(d) builds (default); batch: 256 uneven molecules per build "
                         "(plain FP64 vs batched graph FP64 vs batched graph mixed); serve: Poisson SCF "
                         "requests through the continuous-batching server; gly10: mixed-precision SCF error; "
                         "uhf: open-shell build of (H2O)32+; quick_tests: QUICK's 69 RHF regression inputs "
                         "against the build times of their saved outputs")
    ap.add_argument("--molecules", type=int, default=256)
    ap.add_argument("--rate", type=float, default=20.0, help="serve: arrivals per second")
    ap.add_argument("--requests", type=int, default=60, help="serve: number of requests")
    ap.add_argument("--autotune", type=int, default=1, help="tune launch shapes per class before timing")
    ap.add_argument("--prim-split", type=int, default=0,
                    help="primitive pairs per task piece (0: automatic, < 0: never split)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus %d is not supported "
                                   "(sharding over several GPUs is a design only)" % a.gpus}))
        sys.exit(2)

    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.engine import FockEngine
    from quick_amd.scf import rhf
    from quick_amd.scf import seeded_density as seeded_density_np

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.workload == "batch":
        return run_batch(a, dev)
    if a.workload == "serve":
        return run_serve(a, dev)
    if a.workload == "gly10":
        return run_gly10(a, dev)
    if a.workload == "uhf":
        return run_uhf(a, dev)
    if a.workload == "quick_tests":
        return run_quick_tests(a, dev)
    if a.copy_probe:
        nbytes = a.copy_probe << 20
        x = torch.ones(nbytes, dtype=torch.uint8, device=dev)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "copies": 3}))
        return
    mol = M.water_cluster(a.nwater)
    bs = B.build_basis(mol, "6-31G*")
    log(f"[bench] (H2O){a.nwater} 6-31G(d): {mol.natom} atoms, {bs.nbasis} functions")
    t0 = time.perf_counter()
    eng_replicas = 8
    eng = FockEngine(bs, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff, precision=a.precision)
    if a.prim_split:
        eng.set_prim_split(a.prim_split)
    eng.prepare()
    t_setup = time.perf_counter() - t0
    st = eng.stats()
    log(f"[bench] setup {t_setup:.2f}s  pairs {st['n_pairs']}  tasks fp64 {st['n_tasks_fp64']} "
        f"fp32 {st['n_tasks_fp32']}")

    # ---- SCF (reported, not part of the throughput number) ---------------
    # plain: full FP64 builds every iteration, DIIS and diagonalisation on the
    # host; device: every matrix on the GPU, incremental builds (QUICK's
    # deltaO) with a full rebuild every 8 iterations and at convergence, in
    # FP64 and in mixed precision
    scf = None
    S, T, V = eng.one_electron()
    if not a.no_scf:
        from quick_amd.scf import rhf_device
        eng.fock2e(seeded_density_np(S, mol.nelec() // 2))          # load the kernels once
        res = rhf(eng, etol=1e-8, dtol=1e-6, one_electron=(S, T, V))
        log(f"[bench] SCF (plain) E = {res.energy:.10f} in {res.iterations} it, {res.wall:.2f}s")
        P_host = res.density
        scf = {"energy_eh": res.energy, "iterations": res.iterations, "converged": bool(res.converged),
               "wall_s": round(res.wall, 3), "fock_ms_mean": round(1e3 * float(np.mean(res.fock_times)), 3)}
        Sd = torch.tensor(S, device=dev)
        torch.linalg.eigh(Sd)                                         # warm the device eigensolver
        torch.cuda.synchronize()
        for prec in ("fp64", "mixed"):
            e2 = FockEngine(bs, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff, precision=prec)
            e2.fock2e(P_host)
            r2 = rhf_device(e2, etol=1e-8, dtol=1e-6, one_electron=(S, T, V))
            log(f"[bench] SCF (device, incremental, {prec}) E = {r2.energy:.10f} in {r2.iterations} it "
                f"({r2.full_builds} full), {r2.wall:.2f}s")
            scf[f"device_{prec}"] = {"energy_eh": r2.energy, "dE_vs_plain_eh": abs(r2.energy - res.energy),
                                     "iterations": r2.iterations, "full_builds": r2.full_builds,
                                     "converged": bool(r2.converged), "wall_s": round(r2.wall, 3),
                                     "fock_diis_ms_mean": round(1e3 * float(np.mean(r2.fock_times)), 3),
                                     "speedup_vs_plain": round(res.wall / r2.wall, 2)}
            e2.close()
    else:
        P_host = seeded_density_np(S, mol.nelec() // 2)

    # ---- timed Fock builds -------------------------------------------------
    n = bs.nbasis
    P = torch.from_numpy(np.ascontiguousarray(P_host)).to(dev)
    F = torch.zeros((n, n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    tuned = None
    if a.autotune:
        from quick_amd.autotune import autotune
        ta = time.perf_counter()
        tuned = autotune(eng, sh, builds=2, make_density=lambda i: P.data_ptr())
        torch.cuda.synchronize()
        log(f"[bench] autotune {time.perf_counter() - ta:.2f}s: " +
            ", ".join(f"{k}:{v[0]}" for k, v in sorted(tuned.items())))
    def timed(fn, steps):
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps, time.perf_counter() - w0

    # eager: one launch per class on the stream
    eager = lambda: eng.fock2e_device(P.data_ptr(), F.data_ptr(), sh)
    for _ in range(a.warmup):
        eager()
    ms_eager, _ = timed(eager, a.steps)
    # the same build replayed from a HIP graph (a one-molecule qf_batch over
    # the same resident P and F buffers: no launch gaps between classes)
    from quick_amd.engine import FockBatch
    gb = FockBatch([eng], replicas=eng_replicas, graph=True, density_ptrs=[P.data_ptr()], fock_ptrs=[F.data_ptr()])
    graph = lambda: gb.build(sh)
    for _ in range(a.warmup):
        graph()
    ms_graph, wall = timed(graph, a.steps)
    G_graph, _ = gb.get(0)
    gb.close()
    ms_per_step = ms_graph
    value = 1e3 / ms_per_step
    log(f"[bench] eager {ms_eager:.3f} ms/build, graph replay {ms_graph:.3f} ms/build")
    # per-launch durations (HIP events around every class launch), separately
    eng.set_timing(True)
    for _ in range(max(3, a.steps // 4)):
        eager()
    torch.cuda.synchronize()
    launch_ms, calls = eng.read_timing()
    eng.set_timing(False)

    # ---- roofline of the dominant kernel -------------------------------------
    flops = eng.launch_flops()          # per launch, density of the last build
    names = eng.launches()
    avg_ms = launch_ms / np.maximum(calls, 1)
    k = int(np.argmax(launch_ms))
    peak = FP32_PEAK_TFLOPS if names[k][1] else FP64_PEAK_TFLOPS
    achieved = flops[k] / (avg_ms[k] * 1e-3) / 1e12
    build_flops = float(np.sum(flops))
    nquart, nprim = eng.launch_counts()
    per_class = [{"class": nm, "fp32": f32, "tasks": nt, "avg_ms": round(float(m), 4),
                  "gflop": round(float(fl) / 1e9, 3), "quartets": int(q), "prim_quartets": int(pq),
                  "tflops": round(float(fl) / (m * 1e-3) / 1e12, 3) if m > 0 else None}
                 for (nm, f32, nt), m, fl, q, pq in zip(names, avg_ms, flops, nquart, nprim)]
    log("[bench] per-class: " + json.dumps(per_class))

    # HBM bytes per launch of that kernel, measured by rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes over this workload (scaled by a copy of known size);
    # reported only while the summary names the same kernel
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_h2o{a.nwater}_631gd_{a.precision}.json")
    kname = names[k][0] + (" fp32" if names[k][1] else " fp64")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        if tj.get("kernel") == kname:
            traffic = tj["bytes_per_launch"]

    # ---- CPU baseline: the oracle, one full build on the host cores ---------
    cpu = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        threads = min(16, os.cpu_count() or 1)
        qb = O.QuickBasisRef(mol.symbols, mol.xyz, "6-31G*")
        orc = O.Oracle(qb)
        times = []
        for _ in range(max(1, a.cpu_builds)):
            c0 = time.perf_counter()
            G_cpu, nq = orc.fock2e(P_host, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff, threads=threads)
            times.append(time.perf_counter() - c0)
            log(f"[bench] CPU oracle build {times[-1]:.2f}s on {threads} threads")
        tcpu = float(np.median(times))
        dF = float(np.max(np.abs(G_graph - G_cpu)))
        # one thread: the serial restatement over the first eighth of the bra
        # shells, extrapolated to the whole build by its share of quartets
        nsh = qb.nshell
        c0 = time.perf_counter()
        _, nq1 = orc.fock2e_shell_range(P_host, 0, max(1, nsh // 8), cutoff=a.cutoff, prim_limit=0.1 * a.cutoff)
        t1 = time.perf_counter() - c0
        t1_full = t1 * nq / max(1, nq1)
        log(f"[bench] CPU oracle 1 thread: {nq1} of {nq} quartets in {t1:.2f}s -> {t1_full:.1f}s per build")
        cpu = {"value": round(1.0 / tcpu, 5), "unit": "Fock builds/s", "cores": threads, "kind": "port",
               "sample": f"median of {len(times)} full (H2O){a.nwater} 6-31G(d) builds of the same density "
                         f"({nq} shell quartets each, {tcpu:.2f}s per build, OpenMP over bra shells); "
                         f"max|dF| GPU vs CPU {dF:.1e}",
               "single_thread": {"value": round(1.0 / t1_full, 5), "unit": "Fock builds/s", "cores": 1,
                                 "sample": f"bra shells 0..{max(1, nsh // 8) - 1} of {nsh} ({nq1} of {nq} quartets, "
                                           f"{t1:.2f}s), extrapolated by quartet count"}}

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"(H2O){a.nwater} RHF/6-31G(d) two-electron Fock build G(P), converged density",
                   "natom": mol.natom, "nbasis": n, "integral_cutoff": a.cutoff,
                   "prim_limit": 0.1 * a.cutoff, "precision": a.precision, "seed": M.SEED},
        "roofline": {"bound": "valu", "kernel": kname,
                     "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic},
        "build_tflops": round(build_flops / (ms_per_step * 1e-3) / 1e12, 3),
        "eager_ms_per_step": round(ms_eager, 4),
        "build_gflop": round(build_flops / 1e9, 3),
        "wall_s_timed": round(wall, 3),
        "scf": scf,
        "autotuned": {k: list(v[0]) for k, v in sorted(tuned.items())} if tuned else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def run_batch(a, dev):
    """256 molecules of uneven size (water clusters 1-16, n-alkanes C1-C12),
    6-31G(d): all Fock builds of one SCF iteration.  Plain FP64 (one
    molecule after another) vs one batched build per iteration replayed from
    a HIP graph, in FP64 and in mixed precision."""
    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.autotune import autotune
    from quick_amd.engine import FockBatch, FockEngine
    from quick_amd.scf import seeded_density
    mols = M.mixed_set(a.molecules)
    t0 = time.perf_counter()
    e64, emx, dens = [], [], []
    for m in mols:
        bs = B.build_basis(m, "6-31G*")
        e = FockEngine(bs, cutoff=1e-10)
        S, _, _ = e.one_electron()
        dens.append(seeded_density(S, m.nelec() // 2))
        e64.append(e)
        emx.append(FockEngine(bs, cutoff=1e-10, precision="mixed"))
    nb = [e.n for e in e64]
    log(f"[batch] {len(mols)} molecules, nbasis {min(nb)}..{max(nb)}, setup {time.perf_counter() - t0:.1f}s")
    s = torch.cuda.current_stream(dev).cuda_stream
    # plain FP64: every molecule on its own, as a per-molecule drop-in would
    Pd = [torch.from_numpy(P).to(dev) for P in dens]
    Fd = [torch.zeros_like(p) for p in Pd]

    def plain():
        for e, p, f in zip(e64, Pd, Fd):
            e.fock2e_device(p.data_ptr(), f.data_ptr(), s)
    t_plain = _timed(plain, max(1, a.steps // 4), 1)
    res = {}
    for name, engs in (("fp64", e64), ("mixed", emx)):
        tuned = None
        if a.autotune:
            tuned = autotune(engs, s, builds=1, batch_kw={"replicas": 2}, densities=dens)
        b = FockBatch(engs, replicas=2, graph=True)
        for i, P in enumerate(dens):
            b.set_density(i, P)
        t = _timed(lambda: b.build(s), a.steps, a.warmup)
        res[name] = (t, b, tuned)
    # accuracy of the mixed build against the FP64 build, per molecule
    dF, dE = 0.0, 0.0
    b64, bmx = res["fp64"][1], res["mixed"][1]
    for i, P in enumerate(dens):
        G64, _ = b64.get(i)
        Gmx, _ = bmx.get(i)
        dF = max(dF, float(np.max(np.abs(Gmx - G64))))
        dE = max(dE, 0.5 * abs(float(np.sum(P * (Gmx - G64)))))
    nm = len(mols)
    out = {
        "metric": METRIC,
        "value": round(nm / res["mixed"][0], 3),
        "unit": "Fock builds/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(res["mixed"][0] * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{nm} molecules (water clusters 1-16, n-alkanes C1-C12) RHF/6-31G(d), "
                               "all Fock builds of one SCF iteration per step, batched HIP-graph replay",
                   "nbasis_min": min(nb), "nbasis_max": max(nb), "nbasis_sum": int(sum(nb)),
                   "integral_cutoff": 1e-10, "precision": "mixed (split 1e-4)"},
        "plain_fp64_builds_per_s": round(nm / t_plain, 3),
        "batched_fp64_builds_per_s": round(nm / res["fp64"][0], 3),
        "batched_mixed_builds_per_s": round(nm / res["mixed"][0], 3),
        "speedup_mixed_batched_vs_plain_fp64": round(t_plain / res["mixed"][0], 3),
        "mixed_max_abs_dF_eh": dF, "mixed_max_abs_dE2_eh": dE,
        "autotuned": {k: {n: list(v[0]) for n, v in sorted(r[2].items())} for k, r in res.items() if r[2]},
    }
    for r in res.values():
        r[1].close()
    print(json.dumps(out), flush=True)


def run_serve(a, dev):
    """Poisson SCF requests (molecules of 3-60 atoms from the batch generator,
    basis uniform over 6-31G / 6-31G(d) / def2-SVP, mixed precision) through
    the continuous-batching server: requests/s and p50/p99 latency."""
    from quick_amd.serve import FockServer, request_stream, summarize
    reqs = request_stream(a.requests, a.rate)
    srv = FockServer(dev, precision="mixed")
    # one warm-up request (kernel loading, allocator)
    srv.run(request_stream(1, 1000.0, seed=7))
    results, st = srv.run(reqs, log=None)
    sm = summarize(results, st["wall_s"])
    out = {
        "metric": METRIC, "value": round(sm["requests_per_s"], 4), "unit": "SCF requests/s",
        "n_gpus": 1, "steps": st["steps"], "warmup": 1, "ms_per_step": round(1e3 * st["wall_s"] / max(1, st["steps"]), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{a.requests} Poisson RHF SCF requests at {a.rate}/s, molecules 3-60 atoms, "
                               "basis in {6-31G, 6-31G(d), def2-SVP}, mixed precision, continuous batching",
                   "rate_per_s": a.rate},
        "p50_latency_s": round(sm["p50_s"], 4), "p99_latency_s": round(sm["p99_s"], 4),
        "converged": sm["converged"], "wall_s": round(st["wall_s"], 3),
    }
    print(json.dumps(out), flush=True)


def run_gly10(a, dev):
    """(Gly)10 6-31G(d,p): SCF in FP64 and in mixed precision (FP32 VRR/HRR
    for quartets with Ycut*Ycut < 1e-4 max), dF of one build and dE_SCF."""
    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.engine import FockEngine
    from quick_amd.scf import rhf
    mol = M.polyglycine(10)
    bs = B.build_basis(mol, "6-31G**")
    res = {}
    for prec in ("fp64", "mixed"):
        e = FockEngine(bs, cutoff=1e-11, prim_limit=1e-12, precision=prec)
        S, T, V = e.one_electron()
        r = rhf(e, etol=1e-8, dtol=1e-6, one_electron=(S, T, V))
        P = torch.from_numpy(r.density).to(dev)
        F = torch.zeros_like(P)
        sh = torch.cuda.current_stream(dev).cuda_stream
        t = _timed(lambda: e.fock2e_device(P.data_ptr(), F.data_ptr(), sh), a.steps, a.warmup)
        res[prec] = (r, t, e)
    P64 = res["fp64"][0].density
    G64 = res["fp64"][2].fock2e(P64)
    Gmx = res["mixed"][2].fock2e(P64)
    out = {
        "metric": METRIC, "value": round(1.0 / res["mixed"][1], 3), "unit": "Fock builds/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(res["mixed"][1] * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "(Gly)10 beta strand RHF/6-31G(d,p), split-precision Fock build", "natom": mol.natom,
                   "nbasis": bs.nbasis, "integral_cutoff": 1e-11, "split": 1e-4},
        "fp64_builds_per_s": round(1.0 / res["fp64"][1], 3),
        "scf_energy_fp64": res["fp64"][0].energy, "scf_energy_mixed": res["mixed"][0].energy,
        "dE_scf_eh": abs(res["mixed"][0].energy - res["fp64"][0].energy),
        "max_abs_dF_eh": float(np.max(np.abs(Gmx - G64))),
        "scf_iterations": [res["fp64"][0].iterations, res["mixed"][0].iterations],
    }
    print(json.dumps(out), flush=True)


def run_uhf(a, dev):
    """(H2O)32+ doublet UHF/6-31G(d): the open-shell build (two passes of
    the class kernels: G(Pa+Pb) and the exchange-only G(Pa-Pb)) on the
    converged UHF densities, against the closed-shell build of the same
    basis on Pa+Pb."""
    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.autotune import autotune
    from quick_amd.engine import FockEngine
    from quick_amd.scf import uhf
    m0 = M.water_cluster(a.nwater)
    mol = B.Molecule(m0.symbols, m0.xyz, charge=1)
    bs = B.build_basis(mol, "6-31G*")
    e = FockEngine(bs, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff)
    S, T, V = e.one_electron()
    r = uhf(e, mult=2, etol=1e-8, dtol=1e-6, one_electron=(S, T, V))
    log(f"[bench] UHF E = {r.energy:.10f} in {r.iterations} it, {r.wall:.2f}s")
    Pa, Pb = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in r.density)
    Pt = Pa + Pb
    sh = torch.cuda.current_stream(dev).cuda_stream
    if a.autotune:
        autotune(e, sh, builds=2, make_density=lambda i: Pt.data_ptr())
    Fa, Fb, F = torch.zeros_like(Pa), torch.zeros_like(Pa), torch.zeros_like(Pa)
    t_os = _timed(lambda: e.fock2e_os_device(Pa.data_ptr(), Pb.data_ptr(), Fa.data_ptr(), Fb.data_ptr(), sh),
                  a.steps, a.warmup)
    t_cs = _timed(lambda: e.fock2e_device(Pt.data_ptr(), F.data_ptr(), sh), a.steps, a.warmup)
    out = {
        "metric": METRIC, "value": round(1.0 / t_os, 4), "unit": "open-shell Fock builds/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(t_os * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"(H2O){a.nwater}+ doublet UHF/6-31G(d) open-shell Fock build (Ga, Gb), "
                               "converged UHF densities", "natom": mol.natom, "nbasis": bs.nbasis,
                   "integral_cutoff": a.cutoff, "prim_limit": 0.1 * a.cutoff},
        "closed_shell_ms_same_basis": round(t_cs * 1e3, 4), "open_over_closed": round(t_os / t_cs, 3),
        "uhf_energy_eh": r.energy, "uhf_iterations": r.iterations, "uhf_wall_s": round(r.wall, 3),
    }
    print(json.dumps(out), flush=True)


def run_quick_tests(a, dev):
    """QUICK's own RHF regression inputs (test/*.in, geometries and saved
    outputs in tests/golden/saved_rhf_energies.json): the SCF through the
    host build path (qf_fock_build_host: upload of P, build, download of G,
    as gpu_get_cshell_eri_ does), at each input's cutoff.  Reports the
    median time per Fock build against the saved output's TOTAL 2e TIME per
    SCF cycle (QUICK on an NVIDIA TITAN V, also host wall time around the
    build including its transfers), and the energy against the saved one."""
    from quick_amd import basis as B
    from quick_amd.engine import FockEngine
    from quick_amd.scf import rhf
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "tests", "golden", "saved_rhf_energies.json")) as fh:
        recs = json.load(fh)["records"]
    rows = []
    for r in recs:
        if "t2e_total_s" not in r or not r.get("cycles"):
            continue
        mol = B.Molecule.from_angstrom(r["symbols"], r["xyz_angstrom"], charge=r["charge"])
        try:
            bs = B.build_basis(mol, r["basis"])
        except KeyError:
            continue
        if bs.nbasis != r["nbasis"]:
            continue
        e = FockEngine(bs, cutoff=r["cutoff"], prim_limit=0.1 * r["cutoff"])
        if a.prim_split:
            e.set_prim_split(a.prim_split)
        res = rhf(e, etol=1e-8, dtol=1e-6)
        e.close()
        ours = 1e3 * float(np.median(res.fock_times[1:]))
        quick = 1e3 * r["t2e_total_s"] / r["cycles"]
        rows.append({"name": r["name"], "nbasis": bs.nbasis, "ms_per_build": round(ours, 4),
                     "quick_ms_per_build": round(quick, 4), "speedup": round(quick / ours, 2),
                     "dE_vs_saved_eh": float(abs(res.energy - r["E_total"])), "converged": bool(res.converged)})
        log(f"[bench] {r['name']:28s} n={bs.nbasis:4d} {ours:8.3f} ms/build  QUICK {quick:8.3f}  "
            f"x{quick / ours:6.1f}  dE {abs(res.energy - r['E_total']):.1e}")
    sp = np.array([x["speedup"] for x in rows])
    big = [x for x in rows if x["nbasis"] >= 100]
    out = {
        "metric": "Fock build time vs QUICK's saved TITAN V timings", "value": round(float(np.exp(np.mean(np.log(sp)))), 3),
        "unit": "geometric-mean speedup over QUICK (TITAN V)", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "f64", "data": "QUICK test inputs",
        "config": {"workload": f"{len(rows)} RHF regression inputs of QUICK, host build path incl. transfers",
                   "reference_device": "NVIDIA TITAN V (saved outputs)"},
        "speedup_nbasis_ge_100": [{"name": x["name"], "speedup": x["speedup"]} for x in big],
        "max_dE_vs_saved_eh": max(x["dE_vs_saved_eh"] for x in rows), "rows": rows,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
