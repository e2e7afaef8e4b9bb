"""Denoiser (matrix.world.cpp:952-1037, `ax -n/-N`): the oracle pinned by an
independent route (scipy's correlation with a disc footprint, normalised by
the footprint's overlap with the image) and by the threshold rules, and the
GPU kernel bit-exact against it (same float sums in the same order)."""
import numpy as np
import pytest
from scipy.ndimage import correlate

import ctas_amd as ca
import oracle


def _disc(rad):
    r = np.arange(-rad, rad + 1)
    return (r[:, None] ** 2 + r[None, :] ** 2 <= rad * rad).astype(np.float64)


def _mean(img, rad):
    k = _disc(rad)
    s = correlate(img.astype(np.float64), k, mode="constant", cval=0.0)
    n = correlate(np.ones_like(img, np.float64), k, mode="constant", cval=0.0)
    return s / n


def test_radius_zero_and_constant_image():
    rng = np.random.default_rng(0)
    img = rng.standard_normal((9, 13)).astype(np.float32)
    np.testing.assert_array_equal(oracle.denoise(img, 0, 0.0), img)
    two = np.full((11, 17), 2.0, np.float32)
    np.testing.assert_array_equal(oracle.denoise(two, 3, 0.0), two)


@pytest.mark.parametrize("rad", [1, 2, 3, 5])
def test_mean_matches_disc_correlation(rad):
    rng = np.random.default_rng(rad)
    img = rng.uniform(0.5, 2.0, (23, 31)).astype(np.float32)
    np.testing.assert_allclose(oracle.denoise(img, rad, 0.0), _mean(img, rad), rtol=2e-6)
    # a negative radius is its absolute value (Denoiser stores abs(rad))
    np.testing.assert_array_equal(oracle.denoise(img, -rad, 0.0), oracle.denoise(img, rad, 0.0))


def test_threshold_rules():
    rng = np.random.default_rng(7)
    img = rng.uniform(0.9, 1.1, (20, 20)).astype(np.float32)
    img[5, 7] = 5.0   # outlier
    img[12, 3] = 0.0  # zeros are always replaced
    m = _mean(img, 2)
    out = oracle.denoise(img, 2, 0.5)
    repl = (img == 0) | (np.abs(img - m) > 0.5)
    assert repl[5, 7] and repl[12, 3] and repl.sum() < 10
    np.testing.assert_array_equal(out[~repl], img[~repl])
    np.testing.assert_allclose(out[repl], m[repl], rtol=2e-6)
    out = oracle.denoise(img, 2, -0.3)  # relative to the mean
    repl = (img == 0) | ((m != 0) & (np.abs((img - m) / m) > 0.3))
    np.testing.assert_array_equal(out[~repl], img[~repl])
    np.testing.assert_allclose(out[repl], m[repl], rtol=2e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,rad,thr", [((23, 31), 1, 0.0), ((64, 300), 2, 0.05), ((180, 257), 3, -0.1),
                                           ((7, 5), 5, 0.0), ((1800, 2048), 3, 0.02), ((1, 1), 50, 0.5),
                                           ((1, 7), 2, 0.0), ((2, 2), 5, -0.5), ((3, 1), 0, 0.0)])
def test_gpu_denoise_bit_exact(shape, rad, thr):
    if ca.device_count() < 1:
        pytest.skip("no HIP device")
    rng = np.random.default_rng(rad)
    img = rng.uniform(0.2, 1.0, shape).astype(np.float32)
    img[rng.random(shape) < 0.01] = 0.0
    img[rng.random(shape) < 0.01] *= 4.0
    assert np.array_equal(ca.denoise(img, rad, thr), oracle.denoise(img, rad, thr))
