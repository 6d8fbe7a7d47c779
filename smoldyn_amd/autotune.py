"""Box-size autotuner for batched ensembles.

The virtual-box size (molecules per box, the reference's ``molperbox``
partition) sets the cost balance of a step: small boxes mean more boxes to
scan and more neighbor-box visits, large boxes mean more distance tests per
molecule.  The best value depends on the members' molecule counts, so members
are grouped into size classes (decades of molecule count) and, for each class,
a few candidate values are timed on a short batched run of up to ``probe``
members of that class.  The winner of each class is then applied to all its
members.

Changing the box size changes only which boxes molecules are filed in; each
member stays a valid simulation of its model (the reference lets the user pick
this partition freely), and results of a tuned ensemble equal those of the
same members run alone with the same partition.
"""
import math
import time


def size_class(n):
    """Decade of the molecule count: 1e3-1e4 -> 3, 1e4-1e5 -> 4, ..."""
    return int(math.floor(math.log10(max(n, 1))))


def tune_box_size(make_member, members, counts, candidates=(2.0, 4.0, 8.0), probe=16, steps=30, warmup=5):
    """Pick molecules-per-box per size class.

    make_member(k, molperbox) -> an un-updated Simulation for member k with
    that partition; members: indices; counts[k]: molecule count of member k.
    Returns ({class: best molperbox}, {class: {candidate: ms per step}}).
    """
    from . import Ensemble

    classes = {}
    for k in members:
        classes.setdefault(size_class(counts[k]), []).append(k)
    best, timing = {}, {}
    for c, ks in sorted(classes.items()):
        sample = ks[:probe]
        timing[c] = {}
        for mpb in candidates:
            sims = [make_member(k, mpb) for k in sample]
            ens = Ensemble(sims)
            ens.run_steps(warmup)
            t0 = time.perf_counter()
            ens.run_steps(steps)
            timing[c][mpb] = (time.perf_counter() - t0) * 1e3 / steps
            del ens
        best[c] = min(timing[c], key=timing[c].get)
    return best, timing
