"""The FP32 special functions of the production sweep, evaluated on the device
(thoth_gpu_special_f32) against FP64 references, at their stated bounds.

* Inverse normal of the 32-bit Sobol coordinate x (u = x 2^-32, u = 0 read as
  2^-32): error |z - Phi^-1(u)| / max(|Phi^-1(u)|, 1e-3) <= 3.2e-7 (relative,
  absolute below |z| = 1e-3) against scipy's ndtri, over 4.9M inputs: a random
  sample of all 32-bit inputs, both tails in full (2^18 each), the centre and
  the split w = 8 between the central and the tail polynomial (measured on the
  MI355X: 3.14e-7 at u = 0.709; the fit's FP32 emulation with a correctly
  rounded log gives 3.0e-7); the uniform kernels' split form gives the same
  floats bit for bit.
* erfc (Numerical Recipes' Chebyshev form, fractional error < 1.2e-7 in exact
  arithmetic): relative error against scipy's erfc on the arguments the QE
  exponential branch takes (z / sqrt 2 for every normal the inverse normal can
  return, |x| <= 4.5) at most 2.0e-6 in FP32 on the device (measured 1.87e-6
  at x = 4.37, where erfc = 3e-9: the FP32 exponential of -x^2 + P(t)).
"""
import numpy as np
import pytest
from scipy.special import erfc, ndtri

pytestmark = pytest.mark.gpu

NORMAL_BOUND = 3.2e-7
ERFC_BOUND = 2.0e-6


def _normal_inputs():
    rng = np.random.default_rng(20261204)
    parts = [
        rng.integers(0, 2 ** 32, 1 << 22, dtype=np.uint64),            # all of [0, 2^32)
        np.arange(0, 1 << 18, dtype=np.uint64),                         # lower tail in full
        (1 << 32) - 1 - np.arange(0, 1 << 18, dtype=np.uint64),        # upper tail in full
        (1 << 31) + np.arange(-(1 << 16), 1 << 16).astype(np.int64).astype(np.uint64),  # centre
        np.array([0, 1, 2, (1 << 31) - 1, 1 << 31, (1 << 31) + 1, (1 << 32) - 2, (1 << 32) - 1], np.uint64),
    ]
    # w = -log(4u(1-u)) = 8: the switch from the central to the tail polynomial
    u8 = (1.0 - np.sqrt(1.0 - np.exp(-8.0))) / 2.0
    x8 = int(round(u8 * 2.0 ** 32))
    parts.append(np.arange(x8 - 4096, x8 + 4096, dtype=np.uint64))
    parts.append((1 << 32) - np.arange(x8 - 4096, x8 + 4096, dtype=np.uint64))
    return np.concatenate(parts).astype(np.uint32)


def test_inverse_normal_f32_within_stated_bound(gpu):
    x = _normal_inputs()
    z = gpu.special_f32(0, x).astype(np.float64)
    u = np.maximum(x.astype(np.float64), 1.0) / 2.0 ** 32
    ref = ndtri(u)
    err = np.abs(z - ref) / np.maximum(np.abs(ref), 1e-3)
    worst = int(np.argmax(err))
    print("inverse normal: max error %.3e at x = %d (u = %.6e)" % (err[worst], x[worst], u[worst]))
    assert err.max() <= NORMAL_BOUND
    # symmetry of the integer form: x and 2^32 - x give opposite normals
    inner = x[(x > 0)]
    zi = gpu.special_f32(0, inner)
    zm = gpu.special_f32(0, (np.uint64(1 << 32) - inner.astype(np.uint64)).astype(np.uint32))
    assert np.array_equal(zi, -zm)


def test_split_form_equals_generic_bit_for_bit(gpu):
    x = _normal_inputs()
    assert np.array_equal(gpu.special_f32(0, x).view(np.uint32), gpu.special_f32(1, x).view(np.uint32))


def test_erfc_f32_within_stated_bound(gpu):
    xs = np.concatenate([np.linspace(-4.5, 4.5, 1 << 20), np.array([0.0, -0.0, 1e-30, -1e-30])]).astype(np.float32)
    r = gpu.special_f32(2, xs.view(np.uint32)).astype(np.float64)
    ref = erfc(xs.astype(np.float64))
    err = np.abs(r - ref) / ref
    worst = int(np.argmax(err))
    print("erfc: max relative error %.3e at x = %.6f" % (err[worst], xs[worst]))
    assert err.max() <= ERFC_BOUND
    # the exponential branch reads 1 - u = erfc(z / sqrt 2) / 2 for every normal the sweep draws
    z = gpu.special_f32(0, _normal_inputs())
    arg = (z.astype(np.float64) * 0.70710678118654752).astype(np.float32)
    r = gpu.special_f32(2, arg.view(np.uint32)).astype(np.float64)
    ref = erfc(arg.astype(np.float64))
    assert (np.abs(r - ref) / ref).max() <= ERFC_BOUND
