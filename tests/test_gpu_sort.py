"""GPU: the sort primitives against host sorting, with the reference's own
cases (crates/strata-gpu/src/sort.rs:608-692: sizes 0..4097 over the full
64-bit range, duplicates-heavy unique, pair dedup, per-key minimum) and
sizes past the autotuner's smallest class, under every radix candidate
configuration (digit width x items per thread) and under the tuner."""
from collections import OrderedDict

import numpy as np
import pytest

import strata_k as sk

pytestmark = pytest.mark.gpu


class Lcg:
    """The reference tests' generator (sort.rs:612-621)."""

    def __init__(self, s):
        self.s = s

    def next(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return self.s >> 11


SIZES = [0, 1, 2, 3, 5, 7, 8, 100, 1000, 4097]


@pytest.fixture(params=[-1] + list(range(sk.radix_candidates())), ids=lambda c: f"cand{c}")
def radix_ctx(ctx, request):
    ctx.force_radix(request.param)
    yield ctx
    ctx.force_radix(-1)


def test_sorts_like_host(radix_ctx):
    rng = Lcg(0xA11CE)
    for n in SIZES:
        keys = np.array([rng.next() for _ in range(n)], dtype=np.uint64)
        assert (sk.sort_u64(keys, radix_ctx) == np.sort(keys)).all(), n
    r = np.random.default_rng(5)
    for n in (70_000, 300_001):
        keys = r.integers(0, 2**64, size=n, dtype=np.uint64)
        assert (sk.sort_u64(keys, radix_ctx) == np.sort(keys)).all(), n
        few = r.integers(0, 1 << 13, size=n, dtype=np.uint64)  # two passes of 7-bit digits
        assert (sk.sort_u64(few, radix_ctx) == np.sort(few)).all(), n


def test_many_tiles_and_weights(radix_ctx):
    """5M records (over a thousand tiles and thousands of blocks' digit
    counts in the offset scan) with duplicate-heavy keys and weights: the
    per-key minimum equals the host's."""
    r = np.random.default_rng(55)
    n = 5_000_000
    a = r.integers(0, 1 << 20, size=n, dtype=np.uint32)
    b = r.integers(0, 64, size=n, dtype=np.uint32)
    w = r.integers(-10**9, 10**9, size=n, dtype=np.int64)
    import ctypes as C
    oa, ob, ow = np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(n, np.int64)
    m = C.c_uint64()
    code = radix_ctx.L.sk_sort_min_pairs(radix_ctx.h, a.ctypes.data, b.ctypes.data, w.ctypes.data, n,
                                         oa.ctypes.data, ob.ctypes.data, ow.ctypes.data, C.byref(m))
    assert code == 0
    k = a.astype(np.uint64) << np.uint64(32) | b.astype(np.uint64)
    order = np.lexsort((w, k))
    ks, ws = k[order], w[order]
    first = np.r_[True, ks[1:] != ks[:-1]]
    u = m.value
    assert u == int(first.sum())
    assert (oa[:u].astype(np.uint64) << np.uint64(32) | ob[:u].astype(np.uint64) == ks[first]).all()
    assert (ow[:u] == ws[first]).all()


def test_wide_keys_every_candidate(radix_ctx):
    """A Trop relation of three 30-bit columns (90-bit, u128 keys) with
    duplicates, normalised on load (device sort, per-key minimum), equals a
    host lexsort."""
    from strata_k.core_ir import TROP, CoreProgram, pred
    r = np.random.default_rng(66)
    n = 400_000
    cols = [r.integers(0, 1 << 30, size=n, dtype=np.uint32) for _ in range(2)]
    cols.append(r.integers(0, 50, size=n, dtype=np.uint32))
    cols[0][: n // 2] = cols[0][n // 2:]  # duplicates across halves
    cols[1][: n // 2] = cols[1][n // 2:]
    cols[2][: n // 2] = cols[2][n // 2:]
    w = r.integers(0, 1 << 40, size=n, dtype=np.int64)
    db = radix_ctx.run_semi_naive(CoreProgram([pred("p", 3, TROP)], [], 1), bulk={"p": (cols, w, None)})
    rel = db.relation("p")
    order = np.lexsort((w, cols[2], cols[1], cols[0]))
    c0, c1, c2, ws = cols[0][order], cols[1][order], cols[2][order], w[order]
    first = np.r_[True, (c0[1:] != c0[:-1]) | (c1[1:] != c1[:-1]) | (c2[1:] != c2[:-1])]
    assert (rel.vals[:, 0] == c0[first]).all() and (rel.vals[:, 1] == c1[first]).all()
    assert (rel.vals[:, 2] == c2[first]).all() and (rel.w == ws[first]).all()


def test_unique_like_host(radix_ctx):
    rng = Lcg(0xBEEF)
    for n in [0, 1, 2, 4, 9, 255, 256, 257, 1000, 5000, 90_000]:
        keys = np.array([rng.next() % 40 for _ in range(n)], dtype=np.uint64)
        assert (sk.sort_unique_u64(keys, radix_ctx) == np.unique(keys)).all(), n


def test_pairs_roundtrip(ctx):
    rng = Lcg(0xF00D)
    for _ in range(20):
        m = rng.next() % 500
        pairs = [(rng.next() % 30, rng.next() % 30) for _ in range(m)]
        assert sk.sort_unique_pairs(pairs, ctx) == sorted(set(pairs))


def test_min_by_key_like_host(radix_ctx):
    rng = Lcg(0x5EED)
    for _ in range(30):
        m = rng.next() % 600
        triples = [(rng.next() % 20, rng.next() % 20, rng.next() % 100 - 20) for _ in range(m)]
        want = OrderedDict()
        for s, d, w in sorted(triples):
            want.setdefault((s, d), w)
        assert sk.sort_min_pairs(triples, radix_ctx) == [(s, d, w) for (s, d), w in want.items()]
    r = np.random.default_rng(9)
    n = 200_000
    a, b = r.integers(0, 300, n), r.integers(0, 300, n)
    w = r.integers(-(2**62), 2**62, n)
    got = sk.sort_min_pairs(np.stack([a, b, w], 1), radix_ctx)
    order = np.lexsort((w, b, a))
    a2, b2, w2 = a[order], b[order], w[order]
    first = np.ones(n, bool)
    first[1:] = (a2[1:] != a2[:-1]) | (b2[1:] != b2[:-1])
    assert got == list(zip(a2[first].tolist(), b2[first].tolist(), w2[first].tolist()))


def test_tuner_picks_a_candidate(ctx):
    """Timing every class yields a configuration for every key width; a sort
    of the configuration's own class then equals the host's."""
    ctx.set_autotune(2)
    try:
        for wide in (False, True):
            for weighted in (False, True):
                c = ctx.radix_choice(1 << 20, 40, wide=wide, weighted=weighted)
                desc = sk.radix_candidate(c["candidate"], wide)
                assert 0 <= c["candidate"] < sk.radix_candidates(), c
                assert (c["digit_bits"], c["items_per_thread"], c["block"], c["sweep"]) == (
                    desc["digit_bits"], desc["items_per_thread"], desc["block"], desc["sweep"]), c
                assert c["est_ms"] > 0, c
        keys = np.random.default_rng(3).integers(0, 1 << 40, size=1 << 20, dtype=np.uint64)
        assert (sk.sort_u64(keys, ctx) == np.sort(keys)).all()
    finally:
        ctx.set_autotune(1)
