"""CPU model of the integer map algebra behind the exact cumulative sum.

csrc/measure.hip turns "s <- fl(s + p)" inside one binade into the integer
map m -> m + c[m & 1] (m = s / ulp). These tests restate that algebra in
Python (element map, composition, start/end checks) and check it against
plain sequential IEEE additions, including exact ties, subnormals and zero --
the cases where a wrong rounding rule would flip a sampled index.
"""

import math
import struct

import numpy as np
import pytest

TWO53 = 1 << 53
INVALID = 1 << 60


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def from_bits(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def binade_of(s: float) -> int:
    e = (bits(s) >> 52) & 0x7FF
    return max(e, 1)


def in_binade(s: float, eb: int) -> bool:
    e = (bits(s) >> 52) & 0x7FF
    return e == eb or (eb == 1 and e == 0)


def m_of(s: float, eb: int) -> int:
    return bits(s) - ((eb - 1) << 52)


def from_m(m: int, eb: int) -> float:
    return from_bits(((eb - 1) << 52) + m)


def elem_map(p: float, eb: int):
    b = bits(p)
    ep = (b >> 52) & 0x7FF
    mp = b & ((1 << 52) - 1)
    if ep != 0:
        mp |= 1 << 52
    if mp == 0:
        return (0, 0)
    sh = eb - max(ep, 1)
    if sh < 0:
        return (INVALID, INVALID)
    if sh == 0:
        return (mp, mp)
    if sh >= 63:
        return (0, 0)
    k = mp >> sh
    rem = mp & ((1 << sh) - 1)
    half = 1 << (sh - 1)
    if rem < half:
        return (k, k)
    if rem > half:
        return (k + 1, k + 1)
    return (k + (k & 1), k + ((k + 1) & 1))


def compose(f, g):
    sat = lambda x: min(x, INVALID)
    c0 = sat(f[0] + (g[1] if f[0] & 1 else g[0]))
    c1 = sat(f[1] + (g[1] if (1 + f[1]) & 1 else g[0]))
    return (c0, c1)


def chain_by_maps(s: float, ps):
    """Returns the map-derived end value, or None when the segment leaves the binade."""
    eb = binade_of(s)
    f = (0, 0)
    for p in ps:
        f = compose(f, elem_map(float(p), eb))
    m = m_of(s, eb)
    mend = m + (f[1] if m & 1 else f[0])
    if not in_binade(s, eb) or mend >= TWO53:
        return None
    return from_m(mend, eb)


def chain_seq(s: float, ps):
    for p in ps:
        s = s + float(p)
    return s


@pytest.mark.parametrize("seed", range(6))
def test_maps_equal_sequential_sums(seed):
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(400):
        s = float(rng.uniform(0.5, 1.0)) * 2.0 ** int(rng.integers(-40, 0))
        scale = s * 2.0 ** -float(rng.integers(10, 70))
        ps = rng.random(int(rng.integers(1, 40))) * scale
        got = chain_by_maps(s, ps)
        if got is None:
            continue
        assert got == chain_seq(s, ps)
        checked += 1
    assert checked > 300


def test_exact_ties_round_to_even():
    # p = 1.5 ulp and 0.5 ulp: ties whose direction depends on the parity of m
    for s in (1.0, 1.0 + 2.0**-52, 1.5, 1.5 + 2.0**-52):
        ulp = 2.0**-52
        for p in (0.5 * ulp, 1.5 * ulp, 2.5 * ulp):
            assert chain_by_maps(s, [p]) == s + p
        ps = [0.5 * ulp] * 7 + [1.5 * ulp] * 5
        assert chain_by_maps(s, ps) == chain_seq(s, ps)


def test_subnormal_and_zero_start():
    tiny = 5e-324
    ps = [tiny] * 10 + [3 * tiny, 0.0, 2 * tiny]
    assert chain_by_maps(0.0, ps) == chain_seq(0.0, ps)
    ps = [2.0**-1030] * 50
    assert chain_by_maps(2.0**-1024, ps) == chain_seq(2.0**-1024, ps)


def test_binade_crossing_is_detected():
    s = 1.0 - 2.0**-40
    ps = [2.0**-41] * 4  # ends above 1.0: a new binade, must not use the old grid
    assert chain_by_maps(s, ps) is None
    assert chain_by_maps(0.5, [0.25, 0.25]) is None  # reaches 1.0 exactly


def test_composition_is_associative():
    rng = np.random.default_rng(3)
    eb = binade_of(0.75)
    for _ in range(200):
        fs = [elem_map(float(x), eb) for x in rng.random(3) * 2.0**-45]
        a = compose(compose(fs[0], fs[1]), fs[2])
        b = compose(fs[0], compose(fs[1], fs[2]))
        assert a == b


def chain_tile_chunks(s: float, ps, chunk: int = 64):
    """Python restatement of the wave-parallel fallback for one tile whose
    sum crosses a power of two (csrc/measure.hip chain_tile_chunks): chunks
    whose partial sums stay in one binade advance by their maps and are
    finished from their exact start; crossing chunks add element by element.
    Returns (cdf, end)."""
    n = len(ps)
    chunks = [ps[i:i + chunk] for i in range(0, n, chunk)]
    asums = [float(np.sum(c)) for c in chunks]
    approx = np.cumsum([0.0] + asums)
    maps = []
    for c, (a, z) in zip(chunks, zip(s + approx[:-1], s + approx[1:])):
        eb = binade_of(a)
        lo, hi = math.ldexp(1.0, eb - 1023), math.ldexp(1.0, eb - 1022)
        margin = 1.0 + 2.0**-20
        safe = (eb == 1 or a >= lo * margin) and (z * margin < hi)
        f = (0, 0)
        if safe:
            for p in c:
                f = compose(f, elem_map(float(p), eb))
            safe = f[0] < TWO53 and f[1] < TWO53
        maps.append((safe, eb, f))
    out = []
    for c, (safe, eb, f) in zip(chunks, maps):
        ok = safe and in_binade(s, eb)
        if ok:
            m = m_of(s, eb)
            m2 = m + (f[1] if m & 1 else f[0])
            ok = m2 < TWO53
        if ok:
            x = s
            for p in c:  # the owning lane finishes from the exact start
                x = x + float(p)
                out.append(x)
            s = from_m(m2, eb)
            assert s == x  # the map's end equals the in-order sum
        else:
            for p in c:
                s = s + float(p)
                out.append(s)
    return out, s


@pytest.mark.parametrize("seed", range(4))
def test_chunked_fallback_equals_sequential(seed):
    rng = np.random.default_rng(seed)
    # a tile that starts near zero and crosses many binades, plus exact ties
    ps = rng.random(4096) * np.geomspace(1e-12, 1e-3, 4096)
    ps[::97] = 2.0**-40
    s0 = [0.0, 1e-300, 2.0**-30, 0.5 - 1e-9][seed]
    got, end = chain_tile_chunks(s0, ps)
    want = np.cumsum(np.concatenate([[s0], ps]))[1:]
    assert np.array_equal(np.array(got), want)
    assert end == want[-1]
