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


# --- the floating-point form of the maps (k_tile_map's bulk arithmetic) ----
# The integers c are held in doubles (exact below 2^53): x = |p| 2^(1075 - Eb),
# c0 = x rounded to nearest with ties to even, c1 with ties to odd; the
# composition tests parity as c/2 != floor(c/2). A value >= 2^53 marks an
# unusable map, like the integer form's saturation.


def elem_map_fp(p: float, eb: int):
    try:
        x = math.ldexp(abs(p), 1075 - eb)
    except OverflowError:  # the device's ldexp gives inf
        x = math.inf
    if math.isinf(x):
        return (math.inf, math.inf)
    c0 = float(round(x))  # Python rounds floats half to even
    fl = float(math.floor(x))
    c1 = (fl + fl + 1.0) - c0 if x - fl == 0.5 else c0
    return (c0, c1)


def odd_fp(c: float) -> bool:
    h = c * 0.5
    return math.isfinite(h) and h != math.floor(h)


def compose_fp(f, g):
    return (f[0] + (g[1] if odd_fp(f[0]) else g[0]), f[1] + (g[1] if odd_fp(1.0 + f[1]) else g[0]))


def to_int_map(f):
    return tuple(int(c) if c < TWO53 else INVALID for c in f)


def _same_usable(fi, ff):
    """The integer and the floating-point maps agree: both unusable (>= 2^53)
    or equal."""
    ii = to_int_map(ff)
    for a, b in zip(fi, ii):
        if a >= TWO53:
            assert b >= TWO53
        else:
            assert a == b


def _probe_values(rng, eb):
    """p around the grid of binade eb: multiples and half-multiples of the
    unit (exact ties), random values spanning sh = 0..70, subnormals, zero,
    and values just above the binade (unusable)."""
    u = 2.0 ** (eb - 1075) if eb > 1 else 2.0**-1074
    vals = [0.0, 5e-324, 2.0**-1074 * 3, u, 0.5 * u, 1.5 * u, 2.5 * u, 3.5 * u, 1e6 * u + 0.5 * u,
            float(2**52) * u, float(2**53 - 1) * u]
    vals += [math.ldexp(1.0, e) for e in (eb - 1022, eb - 1021) if e < 1024]
    for sh in range(0, 72):
        mant = rng.integers(2**52, 2**53, size=6)
        for mp in mant:
            e = eb - sh
            if e >= 1:
                vals.append(from_bits(((e) << 52) | (int(mp) & ((1 << 52) - 1))))
        # exact ties at this shift: mp with the low sh bits = 100..0
        if 1 <= sh <= 52 and eb - sh >= 1:
            mp = (int(rng.integers(2**52, 2**53)) >> sh << sh) | (1 << (sh - 1))
            vals.append(from_bits(((eb - sh) << 52) | (mp & ((1 << 52) - 1))))
    return [v for v in vals if math.isfinite(v)]


@pytest.mark.parametrize("eb", [1, 2, 60, 970, 1000, 1022, 1023, 1024, 1100, 2045])
def test_fp_element_map_equals_integer_form(eb):
    rng = np.random.default_rng(eb)
    for p in _probe_values(rng, eb):
        _same_usable(elem_map(p, eb), elem_map_fp(p, eb))


@pytest.mark.parametrize("eb", [1, 1000, 1022, 1023])
def test_fp_composition_equals_integer_form(eb):
    """16-element compositions (one thread of k_tile_map) in both forms, on
    sequences mixing ties, zeros and random values, and sequences that end
    above 2^53 units (unusable)."""
    rng = np.random.default_rng(100 + eb)
    vals = _probe_values(rng, eb)
    u = 2.0 ** (eb - 1075) if eb > 1 else 2.0**-1074
    for trial in range(300):
        if trial % 3 == 0:
            ps = [float(u * (int(rng.integers(0, 8)) + 0.5 * int(rng.integers(0, 2)))) for _ in range(16)]
        else:
            ps = [vals[int(i)] for i in rng.integers(0, len(vals), size=16)]
        fi, ff = (0, 0), (0.0, 0.0)
        for p in ps:
            fi = compose(fi, elem_map(p, eb))
            ff = compose_fp(ff, elem_map_fp(p, eb))
        _same_usable(fi, ff)


def test_fp_maps_drive_the_exact_chain():
    """End to end: a chain driven by the floating-point form equals the
    sequential IEEE sum wherever the integer form applies."""
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(300):
        s = float(rng.uniform(0.5, 1.0)) * 2.0 ** int(rng.integers(-40, 0))
        ps = rng.random(16) * s * 2.0 ** -float(rng.integers(10, 60))
        eb = binade_of(s)
        ff = (0.0, 0.0)
        for p in ps:
            ff = compose_fp(ff, elem_map_fp(float(p), eb))
        f = to_int_map(ff)
        m = m_of(s, eb)
        mend = m + (f[1] if m & 1 else f[0])
        if mend >= TWO53:
            continue
        assert from_m(mend, eb) == chain_seq(s, ps)
        checked += 1
    assert checked > 250


def chain_tile_rounds(s: float, ps, per: int = 16, nthreads: int = 256):
    """Python restatement of the block-parallel sum of one tile whose partial
    sums cross powers of two (csrc/measure.hip chain_tile_rounds): each round
    composes the remaining elements' maps on the grid of the current sum's
    binade, per thread of ``per`` consecutive elements (floating-point form),
    scans them (integer form), walks every thread's elements from its exact
    start, and stops at the first element whose sum leaves the binade, which
    is then added in IEEE arithmetic. Returns (cdf, end, rounds)."""
    n = len(ps)
    vals = [None] * n
    pos, rounds = 0, 0
    while pos < n:
        rounds += 1
        eb = binade_of(s)
        m0 = m_of(s, eb)
        tmaps = []
        for t in range(nthreads):
            fd = (0.0, 0.0)
            for e in range(max(t * per, pos), min(t * per + per, n)):
                fd = compose_fp(fd, elem_map_fp(float(ps[e]), eb))
            tmaps.append(to_int_map(fd))
        excl, cross = (0, 0), n
        for t in range(nthreads):
            m = m0 + (excl[1] if m0 & 1 else excl[0])
            if m < TWO53:
                for e in range(max(t * per, pos), min(t * per + per, n)):
                    c = elem_map_fp(float(ps[e]), eb)
                    cc = c[1] if m & 1 else c[0]
                    if cc < TWO53 and m + int(cc) < TWO53:
                        m += int(cc)
                        vals[e] = from_m(m, eb)
                    else:
                        cross = min(cross, e)
                        break
            excl = compose(excl, tmaps[t])
        if cross == n:
            s, pos = vals[n - 1], n
        else:
            if cross > pos:
                s = vals[cross - 1]
            s = s + float(ps[cross])
            vals[cross] = s
            pos = cross + 1
    return vals, s, rounds


@pytest.mark.parametrize("case", range(7))
def test_tile_rounds_equal_sequential(case):
    rng = np.random.default_rng(case)
    if case < 4:
        # a tile that starts near zero and crosses many binades, plus exact ties
        ps = rng.random(4096) * np.geomspace(1e-12, 1e-3, 4096)
        ps[::97] = 2.0**-40
        s0 = [0.0, 1e-300, 2.0**-30, 0.5 - 1e-9][case]
    elif case == 4:
        # a crossing on almost every element (the sum doubles each time)
        ps = 2.0 ** np.arange(-60, 200, dtype=np.float64) * 1.25
        s0 = 0.0
    elif case == 5:
        # subnormal elements and zeros from zero, then a jump to the normals
        ps = np.concatenate([np.full(500, 5e-324), np.zeros(100), [2.0**-1000], rng.random(300) * 2.0**-1010])
        s0 = 0.0
    else:
        # start exactly on a power of two, ties at the unit
        ps = np.full(3000, 2.0**-53)
        ps[1::2] = 3 * 2.0**-54
        s0 = 1.0
    got, end, rounds = chain_tile_rounds(s0, ps)
    want = np.cumsum(np.concatenate([[s0], ps]))[1:]
    assert np.array_equal(np.array(got), want)
    assert end == want[-1]
    assert rounds >= 1
