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
QUICK's CPU get2e) running one full build of the same workload with OpenMP on
the host cores.
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
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus %d is not supported "
                                   "(the multi-GPU design is described in DESIGN.md)" % a.gpus}))
        sys.exit(2)

    import torch
    from quick_amd import basis as B
    from quick_amd import molecules as M
    from quick_amd.engine import FockEngine
    from quick_amd.scf import rhf

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
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
    eng = FockEngine(bs, cutoff=a.cutoff, prim_limit=0.1 * a.cutoff, precision=a.precision)
    eng.prepare()
    t_setup = time.perf_counter() - t0
    st = eng.stats()
    log(f"[bench] setup {t_setup:.2f}s  pairs {st['n_pairs']}  tasks fp64 {st['n_tasks_fp64']} "
        f"fp32 {st['n_tasks_fp32']}")

    # ---- SCF (reported, not part of the throughput number) ---------------
    scf = None
    S, T, V = eng.one_electron()
    if not a.no_scf:
        res = rhf(eng, etol=1e-8, dtol=1e-6, one_electron=(S, T, V))
        scf = {"energy_eh": res.energy, "iterations": res.iterations, "converged": bool(res.converged),
               "wall_s": round(res.wall, 3), "fock_ms_mean": round(1e3 * float(np.mean(res.fock_times)), 3)}
        log(f"[bench] SCF E = {res.energy:.10f} in {res.iterations} it, {res.wall:.2f}s")
        P_host = res.density
    else:
        from quick_amd.scf import seeded_density
        P_host = seeded_density(S, mol.nelec() // 2)

    # ---- timed Fock builds -------------------------------------------------
    n = bs.nbasis
    P = torch.from_numpy(np.ascontiguousarray(P_host)).to(dev)
    F = torch.zeros((n, n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    for _ in range(a.warmup):
        eng.fock2e_device(P.data_ptr(), F.data_ptr(), sh)
    torch.cuda.synchronize()
    eng.set_timing(True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(a.steps):
        eng.fock2e_device(P.data_ptr(), F.data_ptr(), sh)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_total = ev0.elapsed_time(ev1)
    launch_ms, calls = eng.read_timing()
    eng.set_timing(False)
    ms_per_step = ms_total / a.steps
    value = 1e3 / ms_per_step

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
        G_gpu = eng.fock2e(P_host)
        dF = float(np.max(np.abs(G_gpu - G_cpu)))
        cpu = {"value": round(1.0 / tcpu, 5), "unit": "Fock builds/s", "cores": threads, "kind": "port",
               "sample": f"median of {len(times)} full (H2O){a.nwater} 6-31G(d) builds of the same density "
                         f"({nq} shell quartets each, {tcpu:.2f}s per build, OpenMP over bra shells); "
                         f"max|dF| GPU vs CPU {dF:.1e}"}

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
        "build_gflop": round(build_flops / 1e9, 3),
        "wall_s_timed": round(wall, 3),
        "scf": scf,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
