"""Benchmark: two-electron Fock builds per second on one MI355X.

Headline (``value``, FP64): RHF/6-31G(d) on a seeded 32-water cluster (96
atoms, 608 Cartesian functions), QUICK's screening at integral cutoff 1e-11
and primitive limit 1e-12 (the values QUICK's adjust_Cutoff ends on near
convergence).  The density is the converged SCF density (SCF run first, its
wall time and energy reported beside the throughput), resident in HBM.  A
step is one full Fock build G(P) (screen, all class kernels, symmetrise,
F += G), replayed from a HIP graph.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--precision fp64|mixed]

Prints one JSON line.  Beside the headline it carries, under ``workloads``,
the other workloads of the metric measured in the same run:
* ``split_precision``: (Gly)10 6-31G(d,p) SCF and builds/s in FP64 and in
  mixed precision;
* ``batched``: 256 uneven molecules, plain FP64 per-molecule builds against
  one batched HIP-graph build in mixed precision, both autotuned;
* ``serving``: Poisson SCF requests through the continuous-batching server,
  requests/s and p50/p99 latency at the stated rate;
* ``open_shell``: (H2O)32+ UHF, open-shell builds against the closed-shell
  build of the same basis.
The headline object also carries ``gradient``: the two-electron nuclear
gradient of (H2O)32 at the converged density.
The CPU baseline is the oracle (the C restatement of QUICK's CPU get2e) on
the host cores this job has; the same leg states the GPU results' error in
Eh against that FP64 oracle (Fock matrices and SCF energies).

Single workloads (--workload): batch, serve, gly10, uhf (open-shell build)
and quick_tests (QUICK's own RHF regression inputs against the build times
in their saved outputs).
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


def host_threads():
    """CPU threads this job may use: the affinity mask, capped by
    OMP_NUM_THREADS when the environment sets it (a GPU job's CPU share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


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
    ap.add_argument("--cpu-builds", type=int, default=3, help="oracle builds timed for the CPU baseline (median)")
    ap.add_argument("--workload", default="h2o32", choices=["h2o32", "batch", "serve", "gly10", "uhf", "quick_tests"],
                    help="h2o32: (H2O)32 6-31G(d) builds (default); batch: 256 uneven molecules per build "
                         "(plain FP64 vs batched graph FP64 vs batched graph mixed); serve: Poisson SCF "
                         "requests through the continuous-batching server; gly10: mixed-precision SCF error; "
                         "uhf: open-shell build of (H2O)32+; quick_tests: QUICK's 69 RHF regression inputs "
                         "against the build times of their saved outputs")
    ap.add_argument("--molecules", type=int, default=256)
    ap.add_argument("--rate", type=float, default=20.0, help="serve: arrivals per second")
    ap.add_argument("--requests", type=int, default=60, help="serve: number of requests")
    ap.add_argument("--autotune", type=int, default=1,
                    help="launch shapes per class: 1 = the committed table (quick_amd/tuning/autotune.json) "
                         "when it has this device and workload, else tune; 2 = tune again; 0 = engine defaults")
    ap.add_argument("--tune-out", default=None, help="write the settings tuned in this run to this JSON table")
    ap.add_argument("--prim-split", type=int, default=0,
                    help="primitive pairs per task piece (0: automatic, < 0: never split)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--headline-only", action="store_true",
                    help="headline workload and gradient only (no split-precision, batched, serving and open-shell measurements)")
    ap.add_argument("--serve-rate", type=float, default=10.0, help="default run: serving arrivals per second")
    ap.add_argument("--serve-requests", type=int, default=40, help="default run: serving requests")
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
    out, ctx = measure_h2o32(a, dev)
    ns = None
    if not a.headline_only:
        c3, ctx3 = measure_gly10(a, dev)
        ns = {"split_precision": c3, "batched": measure_batch(a, dev), "serving": measure_serve(a, dev),
              "open_shell": measure_uhf(a, dev)}
        ctx["gly10"] = ctx3
    cpu = None if a.no_cpu_baseline else cpu_baseline_leg(a, ctx)
    if cpu is not None:
        checks = cpu.pop("checks")
        if out["scf"] is not None and "c2_scf_energy_oracle_eh" in checks:
            out["scf"]["oracle_energy_eh"] = checks["c2_scf_energy_oracle_eh"]
            out["scf"]["dE_vs_oracle_eh"] = checks["c2_dE_scf_vs_oracle_eh"]
        if ns is not None and "c3" in checks:
            ns["split_precision"]["vs_fp64_oracle"] = checks["c3"]
        out["vs_fp64_oracle"] = {k[3:]: v for k, v in checks.items() if k.startswith("c2_")}
    out["workloads"] = ns
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def measure_h2o32(a, dev):
    """The headline workload: (H2O)32 6-31G(d) Fock builds on resident
    buffers (graph replay), after the SCF that produces the density."""
    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.engine import FockBatch, FockEngine
    from quick_amd.scf import rhf, rhf_device
    from quick_amd.scf import seeded_density as seeded_density_np
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
        from quick_amd.autotune import tune_or_load
        ta = time.perf_counter()
        tuned, how = tune_or_load(eng, sh, retune=a.autotune == 2, save_path=a.tune_out, tag="h2o32", builds=2,
                                  make_density=lambda i: P.data_ptr())
        torch.cuda.synchronize()
        log(f"[bench] autotune ({how}) {time.perf_counter() - ta:.2f}s: " +
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

    # ---- two-electron nuclear gradient at the same density -------------------
    gd = torch.zeros((mol.natom, 3), dtype=torch.float64, device=dev)
    eng.eri_grad_device(P.data_ptr(), gd.data_ptr(), sh)          # warm-up (allocates its scratch)
    ms_grad, _ = timed(lambda: eng.eri_grad_device(P.data_ptr(), gd.data_ptr(), sh), 3)
    gfl, gq, gpq = eng.grad_launch_counts()
    gradient = {"workload": "two-electron nuclear gradient (QUICK's get_cshell_eri_grad) at the converged density",
                "ms": round(ms_grad, 3), "gflop": round(float(np.sum(gfl)) / 1e9, 3),
                "tflops": round(float(np.sum(gfl)) / (ms_grad * 1e-3) / 1e12, 3),
                "quartets": int(np.sum(gq)), "prim_quartets": int(np.sum(gpq)),
                "over_fock_build": None}
    log(f"[bench] gradient {ms_grad:.2f} ms")

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
        "per_class": per_class,
        "gradient": gradient,
        "autotuned": {k: list(v[0]) for k, v in sorted(tuned.items())} if tuned else None,
    }
    gradient["over_fock_build"] = round(ms_grad / ms_per_step, 2)
    ctx = {"mol": mol, "P": P_host, "G": G_graph, "E_scf": scf["energy_eh"] if scf else None}
    return out, ctx


def measure_gly10(a, dev):
    """Split precision on (Gly)10 6-31G(d,p) (73 atoms, 775 functions): SCF
    with full builds in FP64 and in mixed precision (quartets with Schwarz
    product < 1e-4 max in FP32), then builds/s of both on the FP64-converged
    density, resident buffers, autotuned."""
    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.autotune import tune_or_load
    from quick_amd.engine import FockEngine
    from quick_amd.scf import rhf
    mol = M.polyglycine(10)
    bs = B.build_basis(mol, "6-31G**")
    sh = torch.cuda.current_stream(dev).cuda_stream
    res, engs = {}, {}
    for prec in ("fp64", "mixed"):
        e = FockEngine(bs, cutoff=1e-11, prim_limit=1e-12, precision=prec)
        S, T, V = e.one_electron()
        r = rhf(e, etol=1e-8, dtol=1e-6, one_electron=(S, T, V))
        log(f"[gly10] SCF {prec}: E = {r.energy:.10f} in {r.iterations} it, {r.wall:.2f}s")
        res[prec], engs[prec] = r, e
    P64 = res["fp64"].density
    Pd = torch.from_numpy(np.ascontiguousarray(P64)).to(dev)
    Fd = torch.zeros_like(Pd)
    rate, G = {}, {}
    for prec, e in engs.items():
        if a.autotune:
            tune_or_load(e, sh, retune=a.autotune == 2, save_path=a.tune_out, tag="gly10", builds=2,
                         make_density=lambda i: Pd.data_ptr(), split_modes=(0, 1) if prec == "mixed" else None)
        t = _timed(lambda: e.fock2e_device(Pd.data_ptr(), Fd.data_ptr(), sh), a.steps, a.warmup)
        rate[prec] = 1.0 / t
        G[prec] = e.fock2e(P64)
        log(f"[gly10] {prec}: {1e3 * t:.3f} ms/build")
    st = engs["mixed"].stats()
    out = {"workload": "(Gly)10 beta strand RHF/6-31G(d,p) Fock build, split 1e-4 of max Schwarz product",
           "natom": mol.natom, "nbasis": bs.nbasis, "integral_cutoff": 1e-11,
           "fp64_builds_per_s": round(rate["fp64"], 3), "mixed_builds_per_s": round(rate["mixed"], 3),
           "mixed_over_fp64": round(rate["mixed"] / rate["fp64"], 3),
           "fp32_task_share": round(st["n_tasks_fp32"] / max(1, st["n_tasks_fp32"] + st["n_tasks_fp64"]), 3),
           "scf_energy_fp64_eh": res["fp64"].energy, "scf_energy_mixed_eh": res["mixed"].energy,
           "dE_scf_mixed_vs_fp64_eh": abs(res["mixed"].energy - res["fp64"].energy),
           "max_abs_dF_mixed_vs_fp64_eh": float(np.max(np.abs(G["mixed"] - G["fp64"]))),
           "scf_wall_s": {k: round(v.wall, 3) for k, v in res.items()},
           "scf_iterations": {k: v.iterations for k, v in res.items()}}
    for e in engs.values():
        e.close()
    ctx = {"mol": mol, "P": P64, "G": G, "E": {k: v.energy for k, v in res.items()}}
    return out, ctx


def measure_batch(a, dev):
    """256 uneven molecules (water clusters 1-16, n-alkanes C1-C12),
    6-31G(d): all Fock builds of one SCF iteration.  Plain FP64 (one
    molecule after another, each build on its own) against one batched
    build replayed from a HIP graph, in FP64 and in mixed precision; the
    launch shapes are autotuned on the batch and the plain builds use the
    same engines, so both run under the same tuning."""
    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.autotune import tune_or_load
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
    res = {}
    for name, engs in (("fp64", e64), ("mixed", emx)):
        if a.autotune:
            tune_or_load(engs, s, batched=True, retune=a.autotune == 2, save_path=a.tune_out, tag="mixed256",
                         builds=1, batch_kw={"replicas": 2}, densities=dens,
                         split_modes=(0, 1) if name == "mixed" else None)
        b = FockBatch(engs, replicas=2, graph=True)
        for i, P in enumerate(dens):
            b.set_density(i, P)
        res[name] = (_timed(lambda: b.build(s), a.steps, a.warmup), b)
    # plain FP64: every molecule on its own, as a per-molecule drop-in would,
    # on the FP64 engines (tuned above)
    Pd = [torch.from_numpy(P).to(dev) for P in dens]
    Fd = [torch.zeros_like(p) for p in Pd]

    def plain():
        for e, p, f in zip(e64, Pd, Fd):
            e.fock2e_device(p.data_ptr(), f.data_ptr(), s)
    t_plain = _timed(plain, max(2, a.steps // 4), 1)
    dF, dE = 0.0, 0.0
    b64, bmx = res["fp64"][1], res["mixed"][1]
    for i, P in enumerate(dens):
        G64, _ = b64.get(i)
        Gmx, _ = bmx.get(i)
        dF = max(dF, float(np.max(np.abs(Gmx - G64))))
        dE = max(dE, 0.5 * abs(float(np.sum(P * (Gmx - G64)))))
    for _, b in res.values():
        b.close()
    nm = len(mols)
    log(f"[batch] plain {nm / t_plain:.1f}, batched fp64 {nm / res['fp64'][0]:.1f}, "
        f"batched mixed {nm / res['mixed'][0]:.1f} builds/s")
    return {"workload": f"{nm} molecules (water clusters 1-16, n-alkanes C1-C12) RHF/6-31G(d), all Fock builds "
                        "of one SCF iteration; integral cutoff 1e-10, seeded densities",
            "nbasis_min": min(nb), "nbasis_max": max(nb), "nbasis_sum": int(sum(nb)),
            "plain_fp64_builds_per_s": round(nm / t_plain, 3),
            "batched_fp64_graph_builds_per_s": round(nm / res["fp64"][0], 3),
            "batched_mixed_graph_builds_per_s": round(nm / res["mixed"][0], 3),
            "batched_mixed_over_plain_fp64": round(t_plain / res["mixed"][0], 3),
            "batched_fp64_over_plain_fp64": round(t_plain / res["fp64"][0], 3),
            "mixed_max_abs_dF_vs_fp64_eh": dF, "mixed_max_abs_dE2_vs_fp64_eh": dE}


def measure_serve(a, dev):
    """Poisson SCF requests (molecules of 3-60 atoms from the batch generator,
    basis uniform over 6-31G / 6-31G(d) / def2-SVP, mixed precision) through
    the continuous-batching server."""
    from quick_amd.serve import FockServer, request_stream, summarize
    srv = FockServer(dev, precision="mixed")
    srv.run(request_stream(1, 1000.0, seed=7))          # warm-up request
    results, st = srv.run(request_stream(a.serve_requests, a.serve_rate))
    sm = summarize(results, st["wall_s"])
    log(f"[serve] {sm}")
    return {"workload": f"{a.serve_requests} Poisson RHF SCF requests, molecules 3-60 atoms, basis in "
                        "{6-31G, 6-31G(d), def2-SVP}, mixed precision, continuous batching",
            "rate_per_s": a.serve_rate, "requests_per_s": round(sm["requests_per_s"], 4),
            "p50_latency_s": round(sm["p50_s"], 4), "p99_latency_s": round(sm["p99_s"], 4),
            "converged": sm["converged"], "requests": sm["requests"], "wall_s": round(st["wall_s"], 3),
            "server_steps": st["steps"]}


def cpu_baseline_leg(a, ctx):
    """The CPU baseline: the oracle (C restatement of QUICK's CPU get2e) on
    this job's host threads, timing full builds of the headline workload's
    density (median of --cpu-builds) and one full single-thread build.  The
    same FP64 oracle then states the GPU results' error in Eh: the headline
    Fock matrix, the SCF fixed point (the oracle's SCF from the GPU's
    converged density) and, for the split-precision workload, both GPU Fock
    matrices and both SCF energies."""
    from oracle import oracle as O
    threads = host_threads()
    mol = ctx["mol"]
    P_host = ctx["P"]
    qb = O.QuickBasisRef(mol.symbols, mol.xyz, "6-31G*")
    orc = O.Oracle(qb)
    times = []
    for _ in range(max(1, a.cpu_builds)):
        c0 = time.perf_counter()
        G_cpu, nq = orc.fock2e(P_host, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff, threads=threads)
        times.append(time.perf_counter() - c0)
        log(f"[cpu] oracle build {times[-1]:.2f}s on {threads} threads")
    tcpu = float(np.median(times))
    dF = float(np.max(np.abs(ctx["G"] - G_cpu)))
    c0 = time.perf_counter()
    orc.fock2e(P_host, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff, threads=1)
    t1 = time.perf_counter() - c0
    log(f"[cpu] oracle build {t1:.2f}s on 1 thread")
    checks = {"c2_max_abs_dF_gpu_vs_oracle_eh": dF}
    if ctx.get("E_scf") is not None:
        nel = mol.nelec()
        E_o, _, it = O.rhf(orc, nel, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff, etol=1e-8, dtol=1e-6,
                           threads=threads, guess=P_host)
        checks["c2_scf_energy_oracle_eh"] = float(E_o)
        checks["c2_dE_scf_vs_oracle_eh"] = abs(float(E_o) - ctx["E_scf"])
        log(f"[cpu] oracle SCF from the GPU density: E = {E_o:.10f} ({len(it)} it), dE {checks['c2_dE_scf_vs_oracle_eh']:.1e}")
    g = ctx.get("gly10")
    if g is not None:
        m3 = g["mol"]
        o3 = O.Oracle(O.QuickBasisRef(m3.symbols, m3.xyz, "6-31G**"))
        G3, _ = o3.fock2e(g["P"], cutoff=1e-11, prim_limit=1e-12, threads=threads)
        E3, _, it3 = O.rhf(o3, m3.nelec(), cutoff=1e-11, prim_limit=1e-12, etol=1e-8, dtol=1e-6, threads=threads,
                           guess=g["P"])
        checks["c3"] = {"max_abs_dF_fp64_eh": float(np.max(np.abs(g["G"]["fp64"] - G3))),
                        "max_abs_dF_mixed_eh": float(np.max(np.abs(g["G"]["mixed"] - G3))),
                        "dE2_mixed_eh": 0.5 * abs(float(np.sum(g["P"] * (g["G"]["mixed"] - G3)))),
                        "scf_energy_oracle_eh": float(E3),
                        "dE_scf_fp64_eh": abs(g["E"]["fp64"] - float(E3)),
                        "dE_scf_mixed_eh": abs(g["E"]["mixed"] - float(E3))}
        log(f"[cpu] (Gly)10 vs oracle: {checks['c3']}")
    return {"value": round(1.0 / tcpu, 5), "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": f"median of {len(times)} full (H2O){a.nwater} 6-31G(d) builds of the same density "
                      f"({nq} shell quartets each, {tcpu:.2f}s per build, OpenMP over bra shells)",
            "single_thread": {"value": round(1.0 / t1, 5), "unit": "Fock builds/s", "cores": 1,
                              "sample": f"one full build, {t1:.2f}s"},
            "checks": checks}


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
    from quick_amd.autotune import tune_or_load
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
            tuned, _ = tune_or_load(engs, s, batched=True, retune=a.autotune == 2, save_path=a.tune_out,
                                    tag="mixed256", builds=1, batch_kw={"replicas": 2}, densities=dens,
                                    split_modes=(0, 1) if name == "mixed" else None)
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
    """The split-precision workload alone ((Gly)10 6-31G(d,p), the same
    measurement as the default run's ``workloads.split_precision``, autotuned
    per precision), as the headline line: mixed-precision builds/s."""
    r, _ = measure_gly10(a, dev)
    out = {
        "metric": METRIC, "value": r["mixed_builds_per_s"], "unit": "Fock builds/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 / r["mixed_builds_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {k: r[k] for k in ("workload", "natom", "nbasis", "integral_cutoff")},
    }
    out.update({k: v for k, v in r.items() if k not in out["config"]})
    print(json.dumps(out), flush=True)


def run_uhf(a, dev):
    """The open-shell workload alone, as the headline line."""
    r = measure_uhf(a, dev)
    out = {
        "metric": METRIC, "value": r["builds_per_s"], "unit": "open-shell Fock builds/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {k: r[k] for k in ("workload", "natom", "nbasis", "integral_cutoff", "prim_limit")},
        "closed_shell_ms_same_basis": r["closed_shell_ms"], "open_over_closed": r["open_over_closed"],
        "uhf_energy_eh": r["uhf_energy_eh"], "uhf_iterations": r["uhf_iterations"], "uhf_wall_s": r["uhf_wall_s"],
    }
    print(json.dumps(out), flush=True)


def measure_uhf(a, dev):
    """(H2O)32+ doublet UHF/6-31G(d): the open-shell build (one pass that
    digests every integral with Pa+Pb and, for the exchange terms, with
    Pa-Pb) on the converged UHF densities, against the closed-shell build
    of the same basis on Pa+Pb."""
    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.autotune import tune_or_load
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
    Fa, Fb, F = torch.zeros_like(Pa), torch.zeros_like(Pa), torch.zeros_like(Pa)
    if a.autotune:
        # tuned on the open-shell build itself (its kernels carry more state)
        tune_or_load(e, sh, retune=a.autotune == 2, save_path=a.tune_out, tag="h2o32+", builds=2, os_modes=(0, 1, 2),
                     build=lambda: e.fock2e_os_device(Pa.data_ptr(), Pb.data_ptr(), Fa.data_ptr(), Fb.data_ptr(), sh))
    t_os = _timed(lambda: e.fock2e_os_device(Pa.data_ptr(), Pb.data_ptr(), Fa.data_ptr(), Fb.data_ptr(), sh),
                  a.steps, a.warmup)
    t_cs = _timed(lambda: e.fock2e_device(Pt.data_ptr(), F.data_ptr(), sh), a.steps, a.warmup)
    out = {"workload": f"(H2O){a.nwater}+ doublet UHF/6-31G(d) open-shell Fock build (Ga, Gb), "
                       "converged UHF densities", "natom": mol.natom, "nbasis": bs.nbasis,
           "integral_cutoff": a.cutoff, "prim_limit": 0.1 * a.cutoff,
           "builds_per_s": round(1.0 / t_os, 4), "ms": round(t_os * 1e3, 4),
           "closed_shell_ms": round(t_cs * 1e3, 4), "open_over_closed": round(t_os / t_cs, 3),
           "uhf_energy_eh": r.energy, "uhf_iterations": r.iterations, "uhf_wall_s": round(r.wall, 3)}
    e.close()
    return out


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
