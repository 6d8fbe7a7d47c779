"""Box-size autotuner for batched ensembles.

The virtual-box size (molecules per box, the reference's ``molperbox``
partition) sets the cost balance of a step: small boxes mean more boxes to
scan and more neighbor-box visits, large boxes mean more distance tests per
molecule.  The best value depends on the members' molecule counts, so members
are grouped into size classes (decades of molecule count) and, for each class,
a few candidate values are timed, in interleaved repeats, on batched runs of
up to ``probe`` members of that class.  A candidate replaces the default
(4 molecules per box) only when it is faster beyond the spread of the
repeats; the choice is then applied to all members of the class.

Changing the box size changes only which boxes molecules are filed in; each
member stays a valid simulation of its model (the reference lets the user pick
this partition freely), and results of a tuned ensemble equal those of the
same members run alone with the same partition.
"""
import math
import statistics
import time


def size_class(n):
    """Decade of the molecule count: 1e3-1e4 -> 3, 1e4-1e5 -> 4, ..."""
    return int(math.floor(math.log10(max(n, 1))))


def pick(samples, default):
    """The candidate to use, from repeated timings {candidate: [ms, ...]}.

    A candidate replaces the default only when its slowest repeat is faster
    than the default's fastest one; differences inside the run-to-run spread
    keep the default."""
    fastest = min(samples, key=lambda m: statistics.median(samples[m]))
    if default in samples and max(samples[fastest]) >= min(samples[default]):
        return default
    return fastest


def tune_box_size(make_member, members, counts, candidates=(2.0, 4.0, 8.0), default=4.0, probe=16, steps=100,
                  warmup=10, repeats=3):
    """Pick molecules-per-box per size class.

    make_member(k, molperbox) -> an un-updated Simulation for member k with
    that partition; members: indices; counts[k]: molecule count of member k.
    Every candidate's ensemble is built and warmed up first, then the
    candidates are timed in turn, `repeats` rounds of `steps` steps, so slow
    drifts of the clock or the GPU affect them alike.  Returns
    ({class: chosen molperbox}, {class: {candidate: {"median", "min", "max"}
    in ms per step}}).
    """
    from . import Ensemble

    classes = {}
    for k in members:
        classes.setdefault(size_class(counts[k]), []).append(k)
    best, timing = {}, {}
    for c, ks in sorted(classes.items()):
        sample = ks[:probe]
        ens = {}
        for mpb in candidates:
            ens[mpb] = Ensemble([make_member(k, mpb) for k in sample])
            ens[mpb].run_steps(warmup)
        samples = {mpb: [] for mpb in candidates}
        for _ in range(repeats):
            for mpb in candidates:
                t0 = time.perf_counter()
                ens[mpb].run_steps(steps)
                samples[mpb].append((time.perf_counter() - t0) * 1e3 / steps)
        del ens
        timing[c] = {m: {"median": statistics.median(v), "min": min(v), "max": max(v)} for m, v in samples.items()}
        best[c] = pick(samples, default)
    return best, timing
