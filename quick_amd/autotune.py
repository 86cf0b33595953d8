"""Per-class launch-shape autotuner for the Fock build.

Every class launch of a build is timed with HIP events (launches run one
after another on one stream, so a class's time does not depend on the shape
chosen for the others).  Each candidate shape is applied to all classes at
once, a few builds are timed, and each (bra|ket) class then keeps the
candidate that was fastest for it.  Works on a FockEngine (one molecule) or
on a list of engines that are built as a FockBatch.
"""
from __future__ import annotations

import json
import os

import numpy as np

# (digestion, waves, tasks_per_group, bras_per_task); digestion 0 direct
# atomics, 1 LDS tile, 2 wave per quartet (classes without that kernel use
# the tile; for it tasks_per_group is the ket window); 0 elsewhere = engine
# default
CANDIDATES = [
    (0, 0, 0, 0), (0, 0, 0, 1), (0, 0, 0, 2), (0, 0, 0, 4),
    (1, 4, 16, 0), (1, 4, 8, 0), (1, 2, 8, 0), (1, 1, 4, 0), (1, 4, 32, 0),
    (1, 4, 16, 2), (1, 4, 16, 4), (1, 4, 32, 2), (1, 4, 32, 4),
    (2, 0, 0, 0), (2, 0, 0, 1), (2, 0, 0, 2), (2, 0, 1, 1), (2, 0, 2, 1), (2, 0, 4, 1), (2, 0, 16, 1),
]


# primitive pairs per piece tried after the shape (0: the engine's size by
# molecule, -1: no split)
PIECES = [0, 4, 8, 16, -1]


def _class_times(build, launches, read_timing, set_timing, builds):
    set_timing(True)
    for _ in range(builds):
        build()
    ms, calls = read_timing()
    set_timing(False)
    out = {}
    for (name, f32, _), m, c in zip(launches(), ms, calls):
        out[(name, f32)] = out.get((name, f32), 0.0) + (m / max(c, 1))
    return out


NPC = 10    # pair classes s..f


def _class_index(name):
    names = ["ss", "ps", "pp", "ds", "dp", "dd", "fs", "fp", "fd", "ff"]
    a, b = name.strip("()").split("|")
    return names.index(a), names.index(b)


def autotune(engines, stream, builds=2, candidates=CANDIDATES, batch_kw=None, make_density=None, densities=None,
             pieces=PIECES, build=None, os_modes=None, split_modes=None):
    """Choose a launch shape per class, then a primitive piece size per class.
    ``engines``: one FockEngine or a list (tuned as one FockBatch).  Single
    engine: ``make_density(0)`` returns the device pointer (int) of the
    density to build with, or ``build()`` runs one build of the kind to tune
    (e.g. the open-shell build).  Batch: ``densities`` holds one host density
    per engine.  Returns {class name: (candidate, ms)} (candidate = shape +
    (piece,)) and leaves the best settings applied.  ``os_modes`` (e.g.
    (0, 1, 2) with an open-shell ``build``): a third pass choosing each
    class's open-shell form (fused, two closed-shell passes, fused lean),
    appended to the candidate.  ``split_modes`` (e.g. (0, 1) for mixed
    engines): a last pass choosing per class between FP64 only and the
    precision split, timed over both halves; the candidate then carries
    (..., piece, open-shell form or -1, split mode)."""
    from .engine import FockBatch
    single = not isinstance(engines, (list, tuple))
    engs = [engines] if single else list(engines)

    def timed():
        if single:
            e = engs[0]
            if build is not None:
                return _class_times(build, e.launches, e.read_timing, e.set_timing, builds)
            P = make_density(0)
            return _class_times(lambda: e.fock2e_device(P, 0, stream), e.launches, e.read_timing, e.set_timing,
                                builds)
        b = FockBatch(engs, **(batch_kw or {}))
        for i, P in enumerate(densities):
            b.set_density(i, P)
        t = _class_times(lambda: b.build(stream), b.launches, b.read_timing, b.set_timing, builds)
        b.close()
        return t

    times = {}
    for cand in candidates:
        for e in engs:
            for cb in range(NPC):
                for ck in range(cb + 1):
                    e.set_class_tuning(cb, ck, *cand)
        times[cand] = timed()
    # per class the candidate with the smallest time summed over its FP64 and
    # FP32 launches (they share one shape)
    names = sorted({name for t in times.values() for (name, _) in t})
    chosen = {}
    for name in names:
        def total(c):
            keys = [k for k in times[c] if k[0] == name]
            return sum(times[c][k] for k in keys) if keys else float("inf")
        cand = min(candidates, key=total)
        cb, ck = _class_index(name)
        for e in engs:
            e.set_class_tuning(cb, ck, *cand)
        chosen[name] = (cand, total(cand))
    autotune.last_times = times          # {candidate: {(class, fp32): ms}} of the last call
    # second pass: piece sizes, with every class on its chosen shape
    ptimes = {}
    for pc in pieces or []:
        for e in engs:
            for cb in range(NPC):
                for ck in range(cb + 1):
                    e.set_class_piece(cb, ck, pc)
        ptimes[pc] = timed()
    for name in names:
        cb, ck = _class_index(name)

        def ptotal(pc):
            keys = [k for k in ptimes[pc] if k[0] == name]
            return sum(ptimes[pc][k] for k in keys) if keys else float("inf")
        best = min(pieces, key=ptotal) if pieces else 0
        for e in engs:
            e.set_class_piece(cb, ck, best)
        cand, _ = chosen[name]
        chosen[name] = (tuple(cand) + (best,), ptotal(best) if pieces else chosen[name][1])
    autotune.last_piece_times = ptimes
    if os_modes:
        otimes = {}
        for om in os_modes:
            for e in engs:
                for cb in range(NPC):
                    for ck in range(cb + 1):
                        e.set_class_os(cb, ck, om)
            otimes[om] = timed()
        for name in names:
            cb, ck = _class_index(name)

            def ototal(om):
                keys = [k for k in otimes[om] if k[0] == name]
                return sum(otimes[om][k] for k in keys) if keys else float("inf")
            best = min(os_modes, key=ototal)
            for e in engs:
                e.set_class_os(cb, ck, best)
            cand, _ = chosen[name]
            chosen[name] = (tuple(cand) + (best,), ototal(best))
        autotune.last_os_times = otimes
    if split_modes:
        stimes = {}
        for sm in split_modes:
            for e in engs:
                for cb in range(NPC):
                    for ck in range(cb + 1):
                        e.set_class_split(cb, ck, sm)
            stimes[sm] = timed()
        for name in names:
            cb, ck = _class_index(name)

            def stotal(sm):
                keys = [k for k in stimes[sm] if k[0] == name]
                return sum(stimes[sm][k] for k in keys) if keys else float("inf")
            best = min(split_modes, key=stotal)
            for e in engs:
                e.set_class_split(cb, ck, best)
            cand, _ = chosen[name]
            cand = tuple(cand) + ((-1,) if len(cand) < 6 else ())
            chosen[name] = (cand + (best,), stotal(best))
        autotune.last_split_times = stimes
    return chosen


# ---------------------------------------------------------------------------
# Persisted tables: the chosen settings per class, keyed by device, build
# kind, precision and molecule size, so a run can reuse them instead of
# tuning again.  quick_amd/tuning/autotune.json holds the committed table.
# ---------------------------------------------------------------------------
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "autotune.json")


def tune_grad(engine, grad_call, calls=2):
    """Chooses each class's gradient kernel (per-lane or wave-per-quartet)
    from event-timed gradient calls of both kinds: ``grad_call()`` runs one
    gradient on the engine's stream.  Leaves the faster kernel set per class
    and returns {class name: (mode, ms)}."""
    names = sorted({nm for nm, _, _ in engine.launches()})
    times = {}
    for mode in (1, 0):
        for nm in names:
            engine.set_grad_class(*_class_index(nm), mode)
        grad_call()   # warm-up
        times[mode] = _class_times(grad_call, engine.launches, engine.read_timing, engine.set_timing, calls)
    chosen = {}
    for nm in names:
        tot = {m: sum(v for (k, _), v in times[m].items() if k == nm) for m in (1, 0)}
        best = min(tot, key=tot.get)
        engine.set_grad_class(*_class_index(nm), best)
        chosen[nm] = (best, tot[best])
    tune_grad.last_times = times
    return chosen


def size_bucket(nbasis):
    """The engine's own size classes (primitive piece defaults): <=128, <=256, more."""
    return "n<=128" if nbasis <= 128 else "n<=256" if nbasis <= 256 else "n>256"


def table_key(engines, batched=False, device_name=None, tag=None):
    """device|single or batch|precision|size bucket[|tag]; tag names a
    workload whose settings should not be shared with others of its size."""
    engs = engines if isinstance(engines, (list, tuple)) else [engines]
    if device_name is None:
        try:
            import torch
            props = torch.cuda.get_device_properties(0)
            device_name = getattr(props, "gcnArchName", "").split(":")[0] or props.name
        except Exception:
            device_name = "unknown"
    prec = getattr(engs[0], "precision", "fp64")
    n = max(e.n for e in engs)
    return "|".join([device_name, "batch" if batched else "single", prec, size_bucket(n)] + ([tag] if tag else []))


def load_table(key, path=TABLE):
    """{class name: (digestion, waves, tasks/group, bras/task, piece[, open-shell form[, split mode]])} or None."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        tab = json.load(fh)
    entry = tab.get(key)
    return {k: tuple(v) for k, v in entry.items()} if entry else None


def save_table(key, chosen, path=TABLE):
    """Stores an autotune() result under key (other keys are kept)."""
    tab = {}
    if os.path.exists(path):
        with open(path) as fh:
            tab = json.load(fh)
    tab[key] = {name: [int(x) for x in cand] for name, (cand, _) in sorted(chosen.items())}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as fh:
        json.dump(tab, fh, indent=1, sort_keys=True)


def apply_table(engines, entry):
    engs = engines if isinstance(engines, (list, tuple)) else [engines]
    for name, cand in entry.items():
        cb, ck = _class_index(name)
        for e in engs:
            e.set_class_tuning(cb, ck, *cand[:4])
            e.set_class_piece(cb, ck, cand[4])
            if len(cand) > 5:
                e.set_class_os(cb, ck, cand[5])
            if len(cand) > 6:
                e.set_class_split(cb, ck, cand[6])


def tune_or_load(engines, stream, batched=False, retune=False, path=TABLE, save_path=None, tag=None, **kw):
    """Applies the persisted settings for this key when the table has them
    (returns (entry, "table")), otherwise runs autotune() and stores the
    result in save_path (returns (chosen, "tuned"))."""
    key = table_key(engines, batched, tag=tag)
    if not retune:
        entry = load_table(key, path)
        if entry:
            apply_table(engines, entry)
            return {k: (v, None) for k, v in entry.items()}, "table"
    chosen = autotune(engines, stream, **kw)
    if save_path:
        save_table(key, chosen, save_path)
    return chosen, "tuned"
