"""The oracle's DFT against numpy.fft and a brute-force long-double DFT, with the
FFTW conventions the reference uses (src/DiscreteFourierTransform.cpp:12,22-38):
forward sign -1 unnormalised, inverse sign +1 including the 1/n factor."""
import numpy as np
import pytest

import oracle_binding as ob


def brute_dft(x, sign):
    n = len(x)
    j = np.arange(n)
    ang = sign * 2.0 * np.pi * np.outer(j, j).astype(np.longdouble) / n
    xr = x.real.astype(np.longdouble)
    xi = x.imag.astype(np.longdouble)
    c, s = np.cos(ang), np.sin(ang)
    re = c @ xr - s @ xi
    im = s @ xr + c @ xi
    return (re.astype(np.float64) + 1j * im.astype(np.float64))


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64, 256, 1024])
@pytest.mark.parametrize("variant", [0, 1])
def test_forward_matches_brute_force(n, variant):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = ob.dft(x, -1, variant)
    ref = brute_dft(x, -1)
    assert np.linalg.norm(got - ref) <= 1e-14 * max(np.linalg.norm(ref), 1.0) * np.log2(max(n, 2))


@pytest.mark.parametrize("n", [8, 128, 4096, 65536])
@pytest.mark.parametrize("variant", [0, 1])
def test_forward_inverse_match_numpy(n, variant):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = ob.dft(x, -1, variant)
    assert np.linalg.norm(f - np.fft.fft(x)) <= 1e-14 * np.linalg.norm(f)
    # inverse includes 1/n (Inverse_DFT::compute multiplies by 1/n)
    b = ob.dft(x, +1, variant)
    assert np.linalg.norm(b - np.fft.ifft(x)) <= 1e-14 * np.linalg.norm(b)
    assert np.linalg.norm(ob.dft(f, +1, variant) - x) <= 1e-14 * np.linalg.norm(x)


@pytest.mark.parametrize("n", [3, 12, 100])
def test_non_power_of_two_sizes(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(ob.dft(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    assert np.allclose(ob.dft(x, +1), np.fft.ifft(x), rtol=0, atol=1e-12)


def test_sign_convention_single_tone():
    # x_j = exp(+2 pi i 3 j / n) has its forward spectrum at k = 3 only (sign -1 kernel)
    n = 32
    j = np.arange(n)
    x = np.exp(2j * np.pi * 3 * j / n)
    f = ob.dft(x, -1)
    expect = np.zeros(n, complex)
    expect[3] = n
    assert np.allclose(f, expect, atol=1e-12)


def test_parseval():
    n = 2048
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = ob.dft(x, -1)
    assert abs(np.vdot(f, f).real / n - np.vdot(x, x).real) <= 1e-12 * np.vdot(x, x).real
