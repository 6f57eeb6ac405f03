"""Dihedral columns whose 100-bin modfit histogram has a chosen shape, each with the
bin the reference's scans pick for it (PARENT.cpp:573-607).

With an empty bin the reference walks bins 0 ... 199 (mod 100) and keeps a run of
empty bins only when it is strictly longer than the longest so far: the answer is
the start of the longest run, the first of them to end.  The run through bins 99
and 0 ends last (in the second pass), and the walk's first run, which starts at
bin 0, is then a strict part of it.  With no empty bin the answer is the first of
the least populated bins.  The expected bins below are worked out by hand from
those rules; each case names the rule it exercises.
"""
import numpy as np

NBINS = 100
FRAMES = 400


def _occupied_counts(empty):
    """FRAMES frames over the bins not in `empty`, as evenly as possible."""
    occ = [k for k in range(NBINS) if k not in empty]
    c = np.zeros(NBINS, dtype=np.int64)
    base, extra = divmod(FRAMES, len(occ))
    for i, k in enumerate(occ):
        c[k] = base + (1 if i < extra else 0)
    return c


def _runs(*spans):
    """Bins of the inclusive, possibly wrapping spans (a, b)."""
    out = set()
    for a, b in spans:
        k = a
        while True:
            out.add(k)
            if k == b:
                break
            k = (k + 1) % NBINS
    return out


def _no_empty(low, base=4):
    """Every bin `base` frames except the bins of `low` (bin -> count); the rest
    of FRAMES goes to the first bin that is neither low nor the minimum."""
    c = np.full(NBINS, base, dtype=np.int64)
    for k, n in low.items():
        c[k] = n
    spare = next(k for k in range(NBINS) if k not in low)
    c[spare] += FRAMES - c.sum()
    assert c.min() >= 1
    return c


CASES = [
    # no empty bin: the first of the tied lowest counts
    ("tied minima", _no_empty({17: 1, 60: 1}), 17),
    ("minimum at the last bin", _no_empty({99: 2, 0: 3}), 99),
    ("minimum at bin 0", _no_empty({0: 1, 64: 1}), 0),
    # runs of empty bins
    ("run wrapping through bins 99 and 0 is longest", _occupied_counts(_runs((95, 4), (40, 47))), 95),
    ("two longest runs: the first", _occupied_counts(_runs((20, 29), (60, 69))), 20),
    ("run starting at bin 0, bin 99 occupied", _occupied_counts(_runs((0, 11), (50, 55))), 0),
    ("walk's first run ties another; the wrapping run is longer",
     _occupied_counts(_runs((98, 1), (50, 51))), 98),
    ("walk's first run (bins 0-5) is part of the longest, wrapping run",
     _occupied_counts(_runs((97, 5), (30, 37))), 97),
    ("equal runs, the wrapping one ends last", _occupied_counts(_runs((10, 14), (97, 1))), 10),
    ("one occupied bin (42)", _occupied_counts(set(range(NBINS)) - {42}), 43),
    ("one occupied bin (99)", _occupied_counts(set(range(NBINS)) - {99}), 0),
    ("one occupied bin (0)", _occupied_counts(set(range(NBINS)) - {0}), 1),
    ("one empty bin (99)", _occupied_counts({99}), 99),
    ("one empty bin (0)", _occupied_counts({0}), 0),
    ("one empty bin (63)", _occupied_counts({63}), 63),
    ("one empty bin (64)", _occupied_counts({64}), 64),
    ("run across bins 63 and 64", _occupied_counts(_runs((60, 70), (10, 19))), 60),
    ("run ending at bin 99, bin 0 occupied", _occupied_counts(_runs((90, 99), (5, 13))), 90),
    ("equal runs with one at bin 0", _occupied_counts(_runs((0, 2), (50, 52))), 0),
    ("run from bin 64 wrapping", _occupied_counts(_runs((64, 2), (20, 50))), 64),
]


def bin_width(pi):
    return (2 * pi + 0.000000005) / NBINS


def column(counts, pi, rng):
    """FRAMES dihedral values at the centres of the bins, `counts[k]` in bin k, shuffled."""
    w = bin_width(pi)
    x = np.repeat((np.arange(NBINS) + 0.5) * w, counts)
    assert x.size == FRAMES
    return rng.permutation(x)


def random_counts(rng):
    """A histogram of FRAMES frames over a random subset of bins (sometimes all)."""
    if rng.uniform() < 0.25:
        p = rng.uniform(0.6, 1.0, NBINS)
    elif rng.uniform() < 0.5:
        occ = np.ones(NBINS)
        for _ in range(rng.integers(1, 4)):
            a, n = rng.integers(0, NBINS), rng.integers(1, 60)
            occ[(a + np.arange(n)) % NBINS] = 0
        p = occ * rng.uniform(0.5, 1.0, NBINS)
    else:
        p = (rng.uniform(0, 1, NBINS) > rng.uniform(0.1, 0.9)) * rng.uniform(0.5, 1.0, NBINS)
    if p.sum() == 0:
        p[rng.integers(0, NBINS)] = 1.0
    return rng.multinomial(FRAMES, p / p.sum())
