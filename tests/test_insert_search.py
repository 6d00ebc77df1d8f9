"""kInsert's placement search (kernels_shift.hip, insertSC), emulated lane by
lane on the CPU: lane c in 1..26 holds the number of records its sender
pushes towards the destination with exchange code c; an inclusive scan over
the 64 lanes gives each code's first incoming index; every round of 64
incoming indices then finds, for index k, the largest code whose first index
is at most k with a binary search of five uniform steps that read the other
lanes through shuffles (all 64 lanes take part in every step, so a shuffle
never reads a lane that left the loop). The search must agree with a brute
force over random count patterns, including empty codes, a single busy code
and totals beyond one round."""
import numpy as np
import pytest

N_DIRS = 27


def wave_search(counts):
    """The kernel's arithmetic for one destination: counts[c] for c in 1..26
    (lanes 0 and 27..63 hold zero). Returns the code of every incoming index."""
    cnt = np.zeros(64, np.int64)
    cnt[1:N_DIRS] = counts[1:N_DIRS]
    incl = np.cumsum(cnt)  # waveInclusiveScan
    start = incl - cnt
    total = int(incl[63])
    codes = []
    for k0 in range(0, total, 64):
        k = k0 + np.arange(64)  # one value per lane
        lo = np.ones(64, np.int64)
        hi = np.full(64, N_DIRS - 1, np.int64)
        for _ in range(5):  # ceil(log2(26)) uniform rounds
            mid = (lo + hi + 1) >> 1
            sm = start[mid]  # __shfl(start, mid): every lane is still in the loop
            active = lo < hi
            go = active & (sm <= k)
            lo = np.where(go, mid, lo)
            hi = np.where(active & ~go, mid - 1, hi)
        valid = k < total
        codes.extend(lo[valid].tolist())
    return total, codes


def brute(counts):
    out = []
    for c in range(1, N_DIRS):
        out += [c] * int(counts[c])
    return out


def patterns(seed, n):
    rng = np.random.default_rng(seed)
    for i in range(n):
        counts = np.zeros(N_DIRS, np.int64)
        kind = i % 5
        if kind == 0:  # sparse: most codes empty
            for c in rng.choice(np.arange(1, N_DIRS), rng.integers(0, 5), replace=False):
                counts[c] = rng.integers(1, 40)
        elif kind == 1:  # every code busy
            counts[1:] = rng.integers(1, 20, N_DIRS - 1)
        elif kind == 2:  # one code only, possibly beyond one round
            counts[rng.integers(1, N_DIRS)] = rng.integers(1, 300)
        elif kind == 3:  # first and last codes
            counts[1] = rng.integers(0, 70)
            counts[N_DIRS - 1] = rng.integers(0, 70)
        else:  # random mix with zeros, totals up to several rounds
            counts[1:] = rng.integers(0, 3, N_DIRS - 1) * rng.integers(0, 25, N_DIRS - 1)
        yield counts


@pytest.mark.parametrize("seed", [0, 1])
def test_insert_search_matches_brute_force(seed):
    for counts in patterns(seed, 10000):
        total, codes = wave_search(counts)
        assert total == int(counts[1:].sum())
        assert codes == brute(counts), counts.tolist()


def test_insert_search_all_empty():
    total, codes = wave_search(np.zeros(N_DIRS, np.int64))
    assert total == 0 and codes == []
