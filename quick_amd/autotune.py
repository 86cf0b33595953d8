"""Per-class launch-shape autotuner for the Fock build.

Every class launch of a build is timed with HIP events (launches run one
after another on one stream, so a class's time does not depend on the shape
chosen for the others).  Each candidate shape is applied to all classes at
once, a few builds are timed, and each (bra|ket) class then keeps the
candidate that was fastest for it.  Works on a FockEngine (one molecule) or
on a list of engines that are built as a FockBatch.
"""
from __future__ import annotations

import numpy as np

# (digestion, waves, tasks_per_group, bras_per_task); digestion 0 direct
# atomics, 1 LDS tile, 2 wave per quartet (classes without that kernel use
# the tile); 0 elsewhere = engine default
CANDIDATES = [
    (0, 0, 0, 0), (0, 0, 0, 1), (0, 0, 0, 2), (0, 0, 0, 4),
    (1, 4, 16, 0), (1, 4, 8, 0), (1, 2, 8, 0), (1, 1, 4, 0), (1, 4, 32, 0),
    (1, 4, 16, 2), (1, 4, 16, 4), (1, 4, 32, 2), (1, 4, 32, 4),
    (2, 0, 0, 0), (2, 0, 0, 1), (2, 0, 0, 2),
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
             pieces=PIECES):
    """Choose a launch shape per class, then a primitive piece size per class.
    ``engines``: one FockEngine or a list (tuned as one FockBatch).  Single
    engine: ``make_density(0)`` returns the device pointer (int) of the
    density to build with.  Batch: ``densities`` holds one host density per
    engine.  Returns {class name: (candidate, ms)} (candidate = shape +
    (piece,)) and leaves the best settings applied."""
    from .engine import FockBatch
    single = not isinstance(engines, (list, tuple))
    engs = [engines] if single else list(engines)

    def timed():
        if single:
            e = engs[0]
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
    return chosen
