"""Seeded synthetic inputs for the multi-column sorting workloads.

The reference generates its columns from time(0)-seeded std::rand
(src/generator.h:138-227) and Zipf data with fill_zipf (src/zipf.h:33-69).
Here every column is a pure function of (seed, row): a counter-based
SplitMix64 stream, so host and device see identical inputs and every run
is reproducible.  Zipf group sizes follow fill_zipf's deterministic counts
max(1, floor(n * p_k)) with the rounding remainder handed out one row at a
time from the most frequent rank, so every one of the N values occurs.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def seed_of(config, column):
    """seed = 0xF4C50000 + 16*config + column."""
    return 0xF4C50000 + 16 * config + column


def splitmix64(seed, n, offset=0):
    """Counter-based SplitMix64: value i = mix(seed + (offset+i+1)*golden)."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform_bits(seed, n, bits):
    """Uniform over [0, 2^bits): the top `bits` bits of the stream."""
    z = splitmix64(seed, n)
    return z >> np.uint64(64 - bits) if bits < 64 else z


def uniform_below(seed, n, m):
    """Uniform-ish over [0, m) (z mod m; bias below 2^-24 for m < 2^40)."""
    return splitmix64(seed, n) % np.uint64(m)


def permutation(seed, n):
    """Seeded permutation of range(n)."""
    return np.random.Generator(np.random.PCG64(seed)).permutation(n)


def zipf_counts(n, N, s):
    """fill_zipf counts (src/zipf.h:33-60): max(1, floor(n * p_k)), p_k ∝ k^-s,
    then the remainder n - sum one row each from rank 1 on (or trimmed from
    the largest when the floors of 1 overshoot), so sum == n and all N > 0."""
    k = np.arange(1, N + 1, dtype=np.float64)
    w = 1.0 / np.power(k, s)
    p = w / w.sum()
    c = np.maximum(1, np.floor(n * p)).astype(np.int64)
    diff = n - int(c.sum())
    if diff > 0:
        q, r = divmod(diff, N)
        c += q
        c[:r] += 1
    elif diff < 0:
        c[0] += diff  # the most frequent absorbs the overshoot
        assert c[0] >= 1
    return c


def zipf_column(seed, n, N, s, codes):
    """Column with N distinct values (codes[rank]) at Zipf(s) frequencies,
    rows in a seeded random order."""
    counts = zipf_counts(n, N, s)
    ranks = np.repeat(np.arange(N, dtype=np.int64), counts)
    perm = permutation(seed, n)
    return codes[ranks[perm]]


def bank_dtype(width):
    if width <= 16:
        return np.uint16
    if width <= 32:
        return np.uint32
    return np.uint64


# ---- the five workload shapes -------------------------------------------

def config1(n=1 << 24):
    """ORDER BY c1 (17-bit, 8192 distinct, u32), c2 (33-bit uniform, u64)."""
    codes = permutation(seed_of(1, 9), 1 << 17)[:8192].astype(np.uint32)
    c1 = codes[uniform_below(seed_of(1, 0), n, 8192).astype(np.int64)]
    c2 = uniform_bits(seed_of(1, 1), n, 33).astype(np.uint64)
    return [c1.astype(np.uint32), c2], [17, 33]


def config2(n=1 << 26):
    """Three columns 12/20/30 bits; col0 Zipf(1.0) over 4096 values."""
    codes = permutation(seed_of(2, 9), 4096).astype(np.uint16)
    c0 = zipf_column(seed_of(2, 0), n, 4096, 1.0, codes).astype(np.uint16)
    c1 = uniform_bits(seed_of(2, 1), n, 20).astype(np.uint32)
    c2 = uniform_bits(seed_of(2, 2), n, 30).astype(np.uint32)
    return [c0, c1, c2], [12, 20, 30]


def config3(n=1 << 27, ngroups=1 << 20, s=1.1):
    """c1: 20-bit, 2^20 groups of Zipf(1.1) sizes; c2: 40-bit uniform."""
    bits = max(1, int(np.ceil(np.log2(ngroups))))
    codes = permutation(seed_of(3, 9), 1 << bits)[:ngroups].astype(np.uint32)
    c1 = zipf_column(seed_of(3, 0), n, ngroups, s, codes).astype(np.uint32)
    c2 = uniform_bits(seed_of(3, 1), n, 40).astype(np.uint64)
    return [c1, c2], [20, 40]


def config4(n=1 << 25):
    """Warehouse shape: flag 2b {0,1,2} p=.5/.25/.25, status 1b, date 12b
    over 2526 codes, key 27b uniform."""
    u = uniform_below(seed_of(4, 0), n, 4)
    flag = np.where(u < 2, 0, u - 1).astype(np.uint16)
    status = uniform_bits(seed_of(4, 1), n, 1).astype(np.uint16)
    date = uniform_below(seed_of(4, 2), n, 2526).astype(np.uint16)
    key = uniform_bits(seed_of(4, 3), n, 27).astype(np.uint32)
    return [flag, status, date, key], [2, 1, 12, 27]


def exact_cardinality(col):
    return int(np.unique(col).size)
