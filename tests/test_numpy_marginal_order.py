"""The summation order the marginal-probability kernels reproduce, checked
against numpy itself on the CPU.

np.sum(p, axis=unmeasured) on a C-contiguous (2,)*n array (the reduction of
calculate_probabilities, abstract.py:2739-2741) equals, bit for bit:
* last qubit unmeasured, trailing unmeasured run of L = 2^t elements: per
  output, from 0 and in C order of the outer unmeasured qubits, the pairwise
  sums of its runs in buffers of min(L, 8192) elements;
* last qubit measured: per output, every element added one by one in C order.
The restatement below is the specification of k_marg_np_runs,
k_marg_np_short and k_marg_np_seq (measure.hip); float32 sums stay float32.
"""

import numpy as np
import pytest


def _pairwise(a, dt):
    n = len(a)
    if n < 8:
        res = dt(0)
        for x in a:
            res = dt(res + x)
        return res
    if n <= 128:
        r = [dt(x) for x in a[:8]]
        i = 8
        while i < n - (n % 8):
            for j in range(8):
                r[j] = dt(r[j] + a[i + j])
            i += 8
        res = dt(dt(dt(r[0] + r[1]) + dt(r[2] + r[3])) + dt(dt(r[4] + r[5]) + dt(r[6] + r[7])))
        for x in a[i:]:
            res = dt(res + x)
        return res
    n2 = n // 2
    n2 -= n2 % 8
    return dt(_pairwise(a[:n2], dt) + _pairwise(a[n2:], dt))


def _marginal(p, n, measured, dt, buf=8192):
    ms = sorted(measured)
    out = np.zeros(2 ** len(ms), dtype=dt)
    flat = p.ravel()
    if (n - 1) not in measured:
        t = 0
        while t < n and (n - 1 - t) not in measured:
            t += 1
        L = 2**t
        for idx in range(2 ** (n - t)):
            a = 0
            for q in ms:
                a = (a << 1) | ((idx >> (n - t - 1 - q)) & 1)
            for c0 in range(0, L, buf):
                out[a] = dt(out[a] + _pairwise(flat[idx * L + c0: idx * L + min(L, c0 + buf)], dt))
    else:
        for idx in range(2**n):
            a = 0
            for q in ms:
                a = (a << 1) | ((idx >> (n - 1 - q)) & 1)
            out[a] = dt(out[a] + flat[idx])
    return out


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [3, 6, 9])
def test_marginal_order_equals_numpy(n, dt):
    rng = np.random.default_rng(n)
    p = (rng.random(2**n) ** 3).astype(dt).reshape((2,) * n)
    for mask in range(1, 2**n - 1):
        measured = [q for q in range(n) if (mask >> q) & 1]
        unmeasured = tuple(q for q in range(n) if q not in measured)
        assert np.array_equal(_marginal(p, n, measured, dt), np.sum(p, axis=unmeasured).ravel()), measured


@pytest.mark.parametrize("measured", [[0]])
def test_buffers_of_8192(measured):
    # a trailing run of 2^16 elements is summed buffer after buffer,
    # not as one pairwise tree (with two buffers both orders coincide)
    n = 17
    rng = np.random.default_rng(3)
    p = rng.random(2**n).reshape((2,) * n)
    unmeasured = tuple(q for q in range(n) if q not in measured)
    want = np.sum(p, axis=unmeasured).ravel()
    assert np.array_equal(_marginal(p, n, measured, np.float64), want)
    assert not np.array_equal(_marginal(p, n, measured, np.float64, buf=2**20), want)
