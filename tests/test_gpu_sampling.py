"""Draws without the full cumulative sum (needs an MI355X).

``qsim_sample_exact`` must return, for the same uniforms, exactly the indices
of ``qsim_search_sorted`` over ``qsim_cumsum_exact``'s cdf, which the parity
tests pin bitwise to numpy's ``cumsum`` + ``searchsorted(side="right")`` of
np.random.choice (abstract.py:2774-2781). Checked here on the vectors that
exercise every branch of the chain (maps, binade crossings element by
element, zero runs, spikes, partial last tiles, one element), with uniforms
drawn at random and uniforms placed exactly on normalised cdf values (the
side="right" ties), in float64 and float32, and against numpy directly.
"""

import numpy as np
import pytest
import torch

from qibo_amd import _lib

pytestmark = pytest.mark.gpu


def _vectors(rng):
    out = {}
    p = rng.exponential(size=2**20)
    out["porter-thomas"] = p / p.sum()
    p = rng.exponential(size=2**18 + 77)
    p[rng.random(p.size) < 0.5] = 0.0
    out["sparse-ragged"] = p / p.sum()
    out["geometric"] = 0.5 ** np.arange(1, 1100, dtype=np.float64)
    out["exact-uniform"] = np.full(2**19, 2.0**-19)
    p = rng.random(2**18) * 10.0 ** rng.integers(-300, 0, size=2**18)
    out["wide-range"] = p / p.sum()
    p = np.zeros(50000)
    p[-1] = 1.0
    out["last-only"] = p
    p = np.zeros(9000)
    p[0] = 1.0
    out["first-only"] = p
    p = np.zeros(2**16)
    p[4095], p[4096], p[8191] = 0.25, 0.5, 0.25
    out["tile-edges"] = p
    out["one"] = np.array([1.0])
    p = rng.random(4095)
    out["one-partial-tile"] = p / p.sum()
    p = rng.random(4097)
    out["two-tiles"] = p / p.sum()
    p = rng.random(2**17) ** 8
    out["skewed"] = p / p.sum()
    return out


def _uniforms(p, rng, k=20000):
    u = rng.random(k)
    cdf = np.cumsum(p.astype(np.float64))
    norm = cdf / cdf[-1]
    # exact ties: u equal to normalised cdf values (and their neighbours)
    pick = norm[rng.integers(0, norm.size, size=min(2000, norm.size))]
    pick = pick[pick < 1.0]
    ties = np.concatenate([pick, np.nextafter(pick, 0.0), np.nextafter(pick, 1.0)])
    ties = ties[(ties >= 0.0) & (ties < 1.0)]
    return np.concatenate([u, ties, [0.0, np.nextafter(1.0, 0.0)]])


def _sample(backend, p_dev, u):
    lib = backend.lib
    n, k = p_dev.numel(), len(u)
    ud = torch.as_tensor(u, dtype=torch.float64, device=p_dev.device)
    out = torch.empty(k, dtype=torch.int64, device=p_dev.device)
    stats = torch.empty(3, dtype=torch.float64, device=p_dev.device)
    wsb = lib.qsim_sample_workspace(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=p_dev.device)
    pcode = _lib.F64 if p_dev.dtype == torch.float64 else _lib.F32
    rc = lib.qsim_sample_exact(p_dev.data_ptr(), pcode, n, None, ud.data_ptr(), k, out.data_ptr(), stats.data_ptr(),
                               ws.data_ptr(), wsb, None)
    _lib.check(rc, lib)
    torch.cuda.synchronize()
    return out.cpu().numpy(), stats.cpu().numpy()


def _search_full(backend, p_dev, u):
    cdf = backend.exact_cdf(p_dev)
    return backend.search_cdf(cdf, u).cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sample_exact_matches_full_cdf_search(backend, rng, dtype):
    for name, p in _vectors(rng).items():
        p = p.astype(dtype)
        u = _uniforms(p, rng)
        p_dev = torch.as_tensor(p, device=backend.torch_device)
        got, stats = _sample(backend, p_dev, u)
        want = _search_full(backend, p_dev, u)
        np.testing.assert_array_equal(got, want, err_msg=name)
        # numpy itself: searchsorted on the normalised sequential cumsum
        cdf = np.cumsum(p.astype(np.float64))
        np.testing.assert_array_equal(got, np.searchsorted(cdf / cdf[-1], u, side="right"), err_msg=name)
        assert stats[1] == 0 and stats[2] == 0, name
        assert abs(stats[0] - p.astype(np.float64).sum()) <= 1e-12 * max(1.0, abs(stats[0])), name


def test_sample_exact_invalid_probabilities(backend):
    """NaN, negative and infinite entries are reported in the statistics and
    not sampled (zeros), on vectors long enough that the chain would otherwise
    run element by element."""
    n = 1 << 20
    base = np.full(n, 1.0 / n)
    u = np.linspace(0.0, 0.99, 1000)
    for value, col in ((np.nan, 2), (-1e-9, 1), (np.inf, 0)):
        q = base.copy()
        q[n // 3] = value
        got, stats = _sample(backend, torch.as_tensor(q, device=backend.torch_device), u)
        assert not got.any()
        if col == 0:
            assert np.isinf(stats[0])
        else:
            assert stats[col] == 1


def test_sample_exact_twice_same_workspace(backend, rng):
    """A second call on another vector (smaller, other dtype) in the same
    workspace gives its own draws: nothing of the first call leaks."""
    lib = backend.lib
    p1 = rng.exponential(size=2**20)
    p1 /= p1.sum()
    p2 = rng.random(3000).astype(np.float32)
    p2 /= p2.sum()
    u = rng.random(5000)
    wsb = lib.qsim_sample_workspace(p1.size)
    ws = torch.empty(wsb, dtype=torch.uint8, device=backend.torch_device)
    ud = torch.as_tensor(u, device=backend.torch_device)
    results = []
    for p in (p1, p2):
        t = torch.as_tensor(p, device=backend.torch_device)
        out = torch.empty(u.size, dtype=torch.int64, device=backend.torch_device)
        stats = torch.empty(3, dtype=torch.float64, device=backend.torch_device)
        pcode = _lib.F64 if t.dtype == torch.float64 else _lib.F32
        _lib.check(lib.qsim_sample_exact(t.data_ptr(), pcode, t.numel(), None, ud.data_ptr(), u.size, out.data_ptr(),
                                         stats.data_ptr(), ws.data_ptr(), wsb, None), lib)
        torch.cuda.synchronize()
        cdf = np.cumsum(p.astype(np.float64))
        results.append((out.cpu().numpy(), np.searchsorted(cdf / cdf[-1], u, side="right")))
    for got, want in results:
        np.testing.assert_array_equal(got, want)


def test_sample_exact_argument_errors(backend):
    lib = backend.lib
    assert lib.qsim_sample_workspace(0) == 0
    p = torch.full((16,), 1 / 16, dtype=torch.float64, device=backend.torch_device)
    stats = torch.empty(3, dtype=torch.float64, device=backend.torch_device)
    wsb = lib.qsim_sample_workspace(16)
    ws = torch.empty(wsb, dtype=torch.uint8, device=backend.torch_device)
    assert lib.qsim_sample_exact(p.data_ptr(), _lib.F64, 16, None, None, 0, None, stats.data_ptr(), ws.data_ptr(), wsb,
                                 None) == 0
    assert lib.qsim_sample_exact(p.data_ptr(), _lib.F64, 16, None, None, 5, None, stats.data_ptr(), ws.data_ptr(), wsb,
                                 None) != 0
    assert lib.qsim_sample_exact(p.data_ptr(), _lib.F64, 16, None, None, 0, None, stats.data_ptr(), ws.data_ptr(), wsb - 1,
                                 None) != 0
    assert lib.qsim_sample_exact(p.data_ptr(), 99, 16, None, None, 0, None, stats.data_ptr(), ws.data_ptr(), wsb,
                                 None) != 0
    assert lib.qsim_tile_stats_bytes(0) == 0
    tiles = torch.empty(lib.qsim_tile_stats_bytes(16), dtype=torch.uint8, device=backend.torch_device)
    psi = torch.zeros(16, dtype=torch.complex128, device=backend.torch_device)
    assert lib.qsim_probabilities_tiles(psi.data_ptr(), 4, _lib.C128, p.data_ptr(), tiles.data_ptr(),
                                        tiles.numel() - 1, None) != 0
    assert lib.qsim_probabilities_tiles(psi.data_ptr(), 4, 7, p.data_ptr(), tiles.data_ptr(), tiles.numel(),
                                        None) != 0


@pytest.mark.parametrize("n,dtype", [(0, "complex128"), (11, "complex128"), (12, "complex64"), (21, "complex128"),
                                     (22, "complex64")])
def test_probabilities_with_tile_statistics(backend, backend64, rng, n, dtype):
    """calculate_probabilities of every qubit in order writes p bitwise equal
    to numpy's |psi|^2 and the tile statistics in the same pass; drawing with
    them equals drawing without them and numpy's np.random.choice."""
    be = backend if dtype == "complex128" else backend64
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi = (psi / np.linalg.norm(psi)).astype(dtype)
    p = be.calculate_probabilities(be.cast(psi), list(range(n)), n)
    assert hasattr(p, "_qsim_tiles")
    np.testing.assert_array_equal(p.cpu().numpy(), np.abs(psi) ** 2)
    np.random.seed(77 + n)
    with_tiles = be.sample_shots(p, 30000).cpu().numpy()
    np.random.seed(77 + n)
    plain = be.sample_shots(p.clone(), 30000).cpu().numpy()
    np.random.seed(77 + n)
    want = np.random.choice(np.arange(2**n), size=30000, p=(np.abs(psi) ** 2).astype(np.float64))
    np.testing.assert_array_equal(with_tiles, plain)
    np.testing.assert_array_equal(with_tiles, want)


def test_tile_statistics_not_used_after_an_in_place_change(backend, rng):
    """An in-place change of the probabilities invalidates the statistics
    written with them (the tensor's version counter)."""
    n = 14
    psi = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    psi /= np.linalg.norm(psi)
    p = backend.calculate_probabilities(backend.cast(psi), list(range(n)), n)
    p[: 2**13] = 0.0
    p /= p.sum()
    host = p.cpu().numpy()
    np.random.seed(3)
    got = backend.sample_shots(p, 20000).cpu().numpy()
    assert got.min() >= 2**13
    cdf = np.cumsum(host)
    np.random.seed(3)
    u = np.random.random_sample(20000)
    np.testing.assert_array_equal(got, np.searchsorted(cdf / cdf[-1], u, side="right"))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sample_exact_unaligned_probabilities(backend, rng, dtype):
    """Probabilities that start off a 16-byte boundary (a view one entry in)
    take the element-wise reads and give the same draws as numpy."""
    p = rng.exponential(size=2**16 + 1).astype(dtype)
    p[0] = 0.0
    p /= p.astype(np.float64).sum()
    full = torch.as_tensor(p, device=backend.torch_device)
    view = full[1:]
    assert view.data_ptr() % 16 != 0
    q = p[1:]
    u = _uniforms(q, rng, k=5000)
    got, _ = _sample(backend, view, u)
    cdf = np.cumsum(q.astype(np.float64))
    np.testing.assert_array_equal(got, np.searchsorted(cdf / cdf[-1], u, side="right"))
    np.testing.assert_array_equal(got, _search_full(backend, view.contiguous(), u))
