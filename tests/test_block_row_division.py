"""The block-row division the fetch uses in place of a divide.

For d = xblocks the host stores l = ceil(log2 d) and m = ceil(2^(31+l) / d)
(astc_block.h, set_block_row), and the kernels take the block row of block id
n < 2^31 as mulhi(2n, m) >> l (block_row_of).  The claim is exact division for
every n below 2^31 and every d the encoder can see (1 to 2^22 block columns; the
test goes to 2^31 - 1).  Checked here against integer division: every n for
small d, and for large d the numerators where an error would show first (just
below multiples of d, and the top of the range).
"""
import numpy as np


def magic(d):
    l = (d - 1).bit_length()          # ceil(log2 d), 0 for d = 1
    m = ((1 << (31 + l)) + d - 1) // d
    assert m < (1 << 32)
    return m, l


def quotient(n, m, l):
    prod = (n.astype(np.uint64) << np.uint64(1)) * np.uint64(m)   # < 2^64
    return (prod >> np.uint64(32 + l)).astype(np.int64)


def check(d, n):
    m, l = magic(d)
    n = np.asarray(n, dtype=np.int64)
    n = n[(n >= 0) & (n < (1 << 31))]
    got = quotient(n, m, l)
    want = n // d
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"d={d}: n={n[bad[:3]].tolist()} gives {got[bad[:3]].tolist()}, want {want[bad[:3]].tolist()}"


def test_every_numerator_for_small_divisors():
    n = np.arange(0, 1 << 16, dtype=np.int64)
    for d in range(1, 1025):
        check(d, n)


def test_worst_numerators_for_many_divisors():
    rng = np.random.default_rng(0)
    ds = set(range(1, 5000))
    for k in range(1, 31):
        ds |= {(1 << k) - 1, 1 << k, (1 << k) + 1}
    ds |= set(int(x) for x in rng.integers(1, (1 << 31) - 1, 2000))
    ds |= {4194304, 2796203, 683, 1024, (1 << 31) - 1}
    top = (1 << 31) - 1
    for d in sorted(x for x in ds if x < (1 << 31)):
        q_top = top // d
        qs = np.unique(np.concatenate([np.arange(max(0, q_top - 40), q_top + 1), rng.integers(0, q_top + 1, 40)]))
        n = np.concatenate([qs * d - 1, qs * d, qs * d + d - 1, np.arange(top - 64, top + 1)])
        check(d, n)
