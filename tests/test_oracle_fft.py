"""The oracle's FFT (Stockham mixed radix, symmetric odd-prime DFTs) against numpy.fft and a direct DFT: every length
up to 64 whose prime factors are at most 13 (the reference's FFT sizes are 2^h 3^i 5^j 7^k 11^m,
src/fft/fft_sizes.cpp:10-34) and the lengths of the workloads' grids; a larger prime factor is refused."""

import numpy as np
import pytest

from oracle import oracle

def smooth(n: int) -> bool:
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


LENGTHS = [n for n in range(1, 65) if smooth(n)] + [72, 80, 96, 128, 176, 192, 256, 384, 512, 1000, 2 * 3 * 5 * 7 * 11]


@pytest.mark.parametrize("n", LENGTHS)
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = np.sqrt(n) * np.abs(x).max()
    np.testing.assert_allclose(oracle.fft(x, -1), np.fft.fft(x), rtol=0, atol=4e-15 * scale * max(1.0, np.log2(n)))
    np.testing.assert_allclose(oracle.fft(x, +1), n * np.fft.ifft(x), rtol=0, atol=4e-15 * scale * max(1.0, np.log2(n)))


@pytest.mark.parametrize("n", [6, 11, 48, 91])
def test_fft_matches_direct_sum(n):
    """An independent check of the sign convention: X_k = sum_j x_j exp(sign 2 pi i j k / n), summed in long
    double from exact twiddles."""
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    jk = np.outer(np.arange(n), np.arange(n)) % n
    w = np.exp(-2j * np.pi * jk.astype(np.longdouble) / n)
    np.testing.assert_allclose(oracle.fft(x, -1), w @ x, rtol=0, atol=1e-13 * n)
    np.testing.assert_allclose(oracle.fft(x, +1), np.conj(w) @ x, rtol=0, atol=1e-13 * n)


def test_fft_special_inputs():
    """A delta transforms to a constant, a pure tone to a delta, and Parseval holds."""
    n = 96
    delta = np.zeros(n, complex)
    delta[0] = 1.0
    np.testing.assert_allclose(oracle.fft(delta, -1), np.ones(n), atol=1e-15)
    tone = np.exp(2j * np.pi * 5 * np.arange(n) / n)
    spec = oracle.fft(tone, -1)
    expect = np.zeros(n, complex)
    expect[5] = n
    np.testing.assert_allclose(spec, expect, atol=1e-12)
    x = np.random.default_rng(3).standard_normal(n) + 0j
    assert np.sum(np.abs(oracle.fft(x, -1)) ** 2) == pytest.approx(n * np.sum(np.abs(x) ** 2), rel=1e-14)


@pytest.mark.parametrize("n", [17, 34, 97])
def test_fft_refuses_large_prime_factors(n):
    with pytest.raises(RuntimeError):
        oracle.fft(np.ones(n, complex), -1)
