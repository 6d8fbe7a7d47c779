"""The autotuner's choice rule (smoldyn_amd/autotune.py): a candidate
molecules-per-box value replaces the default only when it wins beyond the
spread of the repeated timings."""
from smoldyn_amd.autotune import pick, size_class


def test_pick_keeps_default_within_noise():
    samples = {2.0: [0.0751, 0.0762, 0.0749], 4.0: [0.0753, 0.0751, 0.0760], 8.0: [0.0800, 0.0810, 0.0805]}
    assert pick(samples, 4.0) == 4.0


def test_pick_takes_clear_winner():
    samples = {2.0: [0.060, 0.061, 0.062], 4.0: [0.075, 0.076, 0.074], 8.0: [0.080, 0.081, 0.079]}
    assert pick(samples, 4.0) == 2.0


def test_pick_without_default_takes_fastest_median():
    assert pick({2.0: [3.0, 1.0, 2.0], 8.0: [2.5, 2.4, 2.6]}, 4.0) == 2.0


def test_size_classes():
    assert [size_class(n) for n in (1, 999, 1000, 99999, 100000)] == [0, 2, 3, 4, 5]
