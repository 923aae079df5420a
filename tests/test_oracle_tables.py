"""Pin the oracle's integer stage (tables, ISE, block layout) without the reference.

Checks, each chosen so a plausible slip fails one of them:
  * trit/quint tables: decoding every table entry with the ASTC specification's
    decoder (written here independently of the oracle's C code) returns the
    tuple; each entry is the largest code of its tuple; the code space is covered.
  * scramble rows: ordering by specification weight unquantisation, whose
    values must be monotone, symmetric (u(i) + u(n-1-i) = 64) and within 1.5 of
    64 i / (n - 1).
  * ISE: every quantisation level round-trips through the oracle's independent
    decoder for random sequences, with the specification's bit count.
  * block mode / assembly: field positions, spec decode of the mode word.
  * the leaf golden blocks are reproduced bit for bit by decode -> re-pack
    (see test_oracle_leaf.py).
"""
import numpy as np
import pytest

from tests import oracle_ref as orc

LEVELS = [2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256]


def bits(v, hi, lo):
    return (v >> lo) & ((1 << (hi - lo + 1)) - 1)


def spec_trits(T):
    if bits(T, 4, 2) == 7:
        C = (bits(T, 7, 5) << 2) | bits(T, 1, 0)
        t4 = t3 = 2
    else:
        C = bits(T, 4, 0)
        if bits(T, 6, 5) == 3:
            t4, t3 = 2, bits(T, 7, 7)
        else:
            t4, t3 = bits(T, 7, 7), bits(T, 6, 5)
    if bits(C, 1, 0) == 3:
        t2, t1, t0 = 2, bits(C, 4, 4), (bits(C, 3, 3) << 1) | (bits(C, 2, 2) & (1 - bits(C, 3, 3)))
    elif bits(C, 3, 2) == 3:
        t2, t1, t0 = 2, 2, bits(C, 1, 0)
    else:
        t2, t1, t0 = bits(C, 4, 4), bits(C, 3, 2), (bits(C, 1, 1) << 1) | (bits(C, 0, 0) & (1 - bits(C, 1, 1)))
    return (t0, t1, t2, t3, t4)


def spec_quints(Q):
    if bits(Q, 2, 1) == 3 and bits(Q, 6, 5) == 0:
        q0b = bits(Q, 0, 0)
        q2 = (q0b << 2) | ((bits(Q, 4, 4) & (1 - q0b)) << 1) | (bits(Q, 3, 3) & (1 - q0b))
        return (4, 4, q2)
    if bits(Q, 2, 1) == 3:
        q2 = 4
        C = (bits(Q, 4, 3) << 3) | ((3 - bits(Q, 6, 5)) << 1) | bits(Q, 0, 0)
    else:
        q2 = bits(Q, 6, 5)
        C = bits(Q, 4, 0)
    if bits(C, 2, 0) == 5:
        return (bits(C, 4, 3), 4, q2)
    return (bits(C, 2, 0), bits(C, 4, 3), q2)


def test_trit_table_inverts_spec_decoder():
    trits, _, _, _ = orc.tables()
    pre = {}
    for code in range(256):
        pre.setdefault(spec_trits(code), []).append(code)
    assert len(pre) == 243
    for idx in range(243):
        t = (idx % 3, idx // 3 % 3, idx // 9 % 3, idx // 27 % 3, idx // 81)
        assert spec_trits(int(trits[idx])) == t
        assert trits[idx] == max(pre[t])


def test_quint_table_inverts_spec_decoder():
    _, quints, _, _ = orc.tables()
    pre = {}
    for code in range(128):
        pre.setdefault(spec_quints(code), []).append(code)
    assert len(pre) == 125
    for idx in range(125):
        q = (idx % 5, idx // 5 % 5, idx // 25)
        assert spec_quints(int(quints[idx])) == q
        assert quints[idx] == max(pre[q])


def spec_unquant_weight(levels, trits, quints, b, sym):
    """Weight unquantisation of the ASTC specification (symbol -> 0..64)."""
    if not trits and not quints:
        acc, n = 0, 0
        while n < 6:
            acc, n = (acc << b) | sym, n + b
        v = acc >> (n - 6)
    elif b == 0:
        return sym * (32 if trits else 16)
    else:
        D, low = sym >> b, sym & ((1 << b) - 1)
        A = 0x7F if low & 1 else 0
        bb, cc = (low >> 1) & 1, (low >> 2) & 1
        if trits:
            C, B = {1: (50, 0), 2: (23, (bb << 6) | (bb << 2) | bb), 3: (11, (cc << 6) | (bb << 5) | (cc << 1) | bb)}[b]
        else:
            C, B = {1: (28, 0), 2: (13, (bb << 6) | (bb << 1) | bb)}[b]
        v = (A & 0x20) | (((D * C + B) ^ A) >> 2)
    return v + 1 if v > 32 else v


def test_scramble_rows_order_spec_unquantisation():
    _, _, scramble, btq = orc.tables()
    for m in range(12):
        b, t, q = btq[m]
        n = LEVELS[m]
        u = [spec_unquant_weight(n, t, q, b, s) for s in range(n)]
        row = [int(x) for x in scramble[m, :n]]
        assert sorted(row) == list(range(n))
        vals = [u[s] for s in row]
        # the decoded weight of linear index i is monotone, symmetric and ~64 i/(n-1)
        assert vals == sorted(vals) and vals[0] == 0 and vals[-1] == 64
        for i in range(n):
            assert vals[i] + vals[n - 1 - i] == 64
            assert abs(vals[i] - 64.0 * i / (n - 1)) <= 1.5
        assert all(x == 0 for x in scramble[m, n:])


def test_bits_trits_quints_table_matches_levels():
    _, _, _, btq = orc.tables()
    for m, n in enumerate(LEVELS):
        b, t, q = btq[m]
        assert (1 << b) * (3 if t else 5 if q else 1) == n


@pytest.mark.parametrize("quant", range(21))
def test_ise_roundtrip_every_level(quant):
    rng = np.random.default_rng(quant)
    _, _, _, btq = orc.tables()
    b, t, q = btq[quant]
    for count in (1, 2, 3, 5, 6, 8, 16):
        nbits_expect = ((8 + 5 * b) * count + 4) // 5 if t else ((7 + 3 * b) * count + 2) // 3 if q else b * count
        if nbits_expect > 128 - 8:
            continue
        for _ in range(20):
            vals = rng.integers(0, LEVELS[quant], size=count).astype(np.uint32)
            words, nbits = orc.bise(vals, quant)
            assert nbits == nbits_expect
            got, nb2 = orc.ise_decode(words, 0, count, quant)
            assert nb2 == nbits_expect
            np.testing.assert_array_equal(got, vals)


def test_ise_exhaustive_trit_and_quint_groups():
    """Every 5-trit group (QUANT_6: 1 bit + trit) and 3-quint group (QUANT_10)."""
    for idx in range(243):
        tr = [idx % 3, idx // 3 % 3, idx // 9 % 3, idx // 27 % 3, idx // 81]
        for low in (0, 1):
            vals = np.array([(x << 1) | low for x in tr], np.uint32)
            words, _ = orc.bise(vals, 4)
            got, _ = orc.ise_decode(words, 0, 5, 4)
            np.testing.assert_array_equal(got, vals)
    for idx in range(125):
        qs = [idx % 5, idx // 5 % 5, idx // 25]
        vals = np.array([(x << 1) | 1 for x in qs], np.uint32)
        words, _ = orc.bise(vals, 6)
        got, _ = orc.ise_decode(words, 0, 3, 6)
        np.testing.assert_array_equal(got, vals)


def test_block_modes_decode_to_4x4_grid():
    assert orc.lib().oracle_blockmode(4) == 0x043
    assert orc.lib().oracle_blockmode(7) == 0x251
    # a block with each mode decodes (spec) to a 4x4 single-plane grid of the right range
    for has_alpha, wq, cem in ((True, 4, 12), (False, 7, 8)):
        blk = orc.pack_block(np.arange(8, dtype=np.uint32) * 30, np.zeros(16, np.uint32), has_alpha)
        ok, _, info = orc.decode_block(blk, 4)
        assert ok == 1 and info.grid_w == 4 and info.grid_h == 4 and info.dual_plane == 0
        assert info.weight_quant == wq and info.cem == cem and info.partitions == 1
        assert info.endpoint_quant == 20


def test_weight_stream_is_bit_reversed_from_bit_127():
    for k in range(64):
        wt = np.zeros(4, np.uint32)
        wt[k // 32] = np.uint32(1) << np.uint32(k % 32)
        out = np.zeros(16, np.uint8)
        orc.lib().oracle_assemble_block(0, 0, orc._ptr(np.zeros(4, np.uint32)), orc._ptr(wt), orc._ptr(out))
        bitpos = 127 - k
        expect = np.zeros(16, np.uint8)
        expect[bitpos // 8] = 1 << (bitpos % 8)
        np.testing.assert_array_equal(out, expect)


def test_endpoint_stream_starts_at_bit_17_and_header_fields():
    ep = np.array([0xFFFFFFFF, 0xFFFFFFFF, 0, 0], np.uint32)
    out = np.zeros(16, np.uint8)
    orc.lib().oracle_assemble_block(0x043, 12, orc._ptr(ep), orc._ptr(np.zeros(4, np.uint32)), orc._ptr(out))
    v = int.from_bytes(out.tobytes(), "little")
    assert v & 0x7FF == 0x043                  # block mode, 11 bits
    assert (v >> 11) & 3 == 0                  # partition count - 1
    assert (v >> 13) & 0xF == 12               # CEM
    assert (v >> 17) & ((1 << 64) - 1) == (1 << 64) - 1
    assert v >> 81 == 0


def test_length_threshold_equivalence():
    """length(v) < 1e-5f  <=>  dot(v, v) < threshold, checked around the threshold."""
    thr = np.float32(orc.lib().oracle_length_threshold())
    bitsv = thr.view(np.uint32)
    xs = (np.arange(-200000, 200000, dtype=np.int64) + int(bitsv)).astype(np.uint32).view(np.float32)
    np.testing.assert_array_equal(np.sqrt(xs) < np.float32(1e-5), xs < thr)
    assert int(bitsv) == 0x2EDBE6FD


def test_newton_rsqrt_accuracy():
    """The canonical schedule's rsqrt stays within 4 ulp of 1/sqrt over the
    range normalize sees (squared lengths from the 1e-10 threshold upward)."""
    f = orc.lib().oracle_rsqrt_nr
    xs = np.float32(10.0) ** np.linspace(-10, 12, 20001).astype(np.float32)
    got = np.array([f(float(x)) for x in xs], np.float32)
    exact = 1.0 / np.sqrt(xs.astype(np.float64))
    ulp = np.spacing(exact.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - exact) / ulp) <= 4.0
