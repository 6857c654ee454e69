"""The oracle's FFT against a direct DFT and against the reference's own
host FFT (runtime/mdrt_fft.c over pocketfft, built into oracle/_ref)."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 16, 35, 54, 64, 97, 216])
@pytest.mark.parametrize("sign", [-1, 1])
def test_line_against_dft(n, sign):
    rng = np.random.default_rng(n)
    data = rng.standard_normal(2 * n)
    ref = np.zeros(2 * n)
    O.lib().orc_dft_line(n, data.ctypes.data_as(O.PD), ref.ctypes.data_as(O.PD), sign)
    got = data.copy()
    O.lib().orc_fft_line(n, got.ctypes.data_as(O.PD), sign)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("K", [(8, 8, 8), (12, 10, 14), (35, 35, 35), (16, 24, 9)])
def test_3d_against_numpy_and_roundtrip(K):
    rng = np.random.default_rng(sum(K))
    real = rng.standard_normal(K)
    spec = O.fft3d_r2c(real)
    assert np.allclose(spec, np.fft.rfftn(real), rtol=0, atol=1e-10 * np.abs(spec).max())
    back = O.fft3d_c2r(spec, K[2])
    assert np.allclose(back / np.prod(K), real, atol=1e-12)


@pytest.mark.skipif(not O.ref_fft_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("K", [(8, 8, 8), (24, 20, 30), (35, 35, 35), (64, 64, 64)])
def test_3d_against_reference_host_fft(K):
    rng = np.random.default_rng(7)
    real = rng.standard_normal(K)
    ours = O.fft3d_r2c(real)
    theirs = O.ref_fft3d_r2c(real)
    assert np.abs(ours - theirs).max() <= 1e-12 * np.abs(theirs).max()
    # backward: the imaginary parts of the zero and Nyquist planes are ignored
    spec = theirs + 1e-3j
    assert np.abs(O.fft3d_c2r(spec, K[2]) - O.ref_fft3d_c2r(spec, K[2])).max() \
        <= 1e-11 * np.abs(real).max() * np.prod(K)
