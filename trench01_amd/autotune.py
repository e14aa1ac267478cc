"""Autotuner for the push schedule.

The push's results do not depend on its schedule (each particle is traced
independently; tallies are integer atomics), but its speed does: queue order
(long-lived electrons first or last), the wave refill threshold, the number of
steps between refill checks, the multi-pass step cap, the workgroups resident
per CU (which also picks the register budget of the ensemble push), and
whether waves carrying old particles get issue priority.  ``tune_push`` times
candidate schedules on a representative batch already resident on the device,
``tune_ensemble`` on one device-injected batch of an ensemble (the same
particles for every candidate); both check that every candidate returns
identical tallies and return the fastest.  Results of ``tune_push`` are
cached per (device, grid, batch size) in a JSON file.
"""
import itertools
import json
import os
from pathlib import Path

DEFAULT_SPACE = dict(order=(0, 1), refill=(16, 32), check_steps=(2, 4, 8), prio_steps=(0,),
                     pass_steps=(0, 150, 250, 400, 600, 1000))
ENSEMBLE_SPACE = dict(pass_steps=(0, 250, 500), refill=(16, 32), check_steps=(4, 8), blocks_per_cu=(0, 4))
CACHE = Path(os.environ.get("TRENCH_HIP_TUNE_CACHE", Path.home() / ".cache" / "trench01_amd" / "push_tune.json"))


def candidates(space=None):
    space = space or DEFAULT_SPACE
    keys = sorted(space)
    for values in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, values))


def apply(prm, sched):
    for k, v in sched.items():
        setattr(prm, k, int(v))
    return prm


def tune_push(tracer, prm, offset, n, space=None, repeats=2, log=None):
    """Time every schedule on pool[offset, offset+n) with the current field.
    Returns (best schedule dict, table of (schedule, best ms))."""
    table = []
    ref = None
    for sched in candidates(space):
        apply(prm, sched)
        best = None
        for _ in range(repeats):
            tally, counts, ms = tracer.trace_resident(offset, n, prm, with_counts=True)
            key = (tally["escaped"], tally["lost"], tally["steps"], counts[0].tobytes(), counts[1].tobytes())
            if ref is None:
                ref = key
            elif key != ref:
                raise RuntimeError(f"schedule {sched} changed the push's results")
            best = ms if best is None else min(best, ms)
        table.append((dict(sched), best))
        if log:
            log(f"  {sched}: {best:.3f} ms")
    table.sort(key=lambda r: r[1])
    return table[0][0], table


def tune_ensemble(ens, prm, space=None, seed=0, repeats=1, log=None):
    """Time every schedule on one device-injected batch of ``ens`` (batch
    index 0 of ``seed`` for every candidate).  Returns (best schedule dict,
    table of (schedule, best ms))."""
    table = []
    ref = None
    for sched in candidates(space or ENSEMBLE_SPACE):
        apply(prm, sched)
        best = None
        for _ in range(repeats):
            ms = ens.run(prm, seed=seed, first_batch=0, batches=1, inject=True)
            counts, scal = ens.totals()
            key = (counts.tobytes(), scal.tobytes())
            if ref is None:
                ref = key
            elif key != ref:
                raise RuntimeError(f"schedule {sched} changed the ensemble's results")
            best = ms if best is None else min(best, ms)
        table.append((dict(sched), best))
        if log:
            log(f"  {sched}: {best:.3f} ms")
    table.sort(key=lambda r: r[1])
    return table[0][0], table


def cache_key(tracer, n):
    nx, ny = tracer.vac.shape
    return f"{tracer.name}|{nx}x{ny}|{n}"


def load_cached(tracer, n):
    try:
        return json.loads(CACHE.read_text()).get(cache_key(tracer, n))
    except (OSError, ValueError):
        return None


def store(tracer, n, sched):
    try:
        data = json.loads(CACHE.read_text()) if CACHE.exists() else {}
    except (OSError, ValueError):
        data = {}
    data[cache_key(tracer, n)] = sched
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        CACHE.write_text(json.dumps(data, indent=1))
    except OSError:
        pass
