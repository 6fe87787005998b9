"""The device's half-angle sin / cos (rk_device.h rk_sincos): Cody-Waite reduction by pi / 2 in
three parts and the Cephes single-precision minimax polynomials, restated here in float32 (each
fused multiply-add rounded once) and checked against float64 sin / cos: a wrong constant, sign or
quadrant fails these bounds.  The device code itself is checked through FK against the FP64
oracle and the reference's ground-truth poses (tests/test_gpu_kinematics.py)."""
import numpy as np

F = np.float32


def _fma(a, b, c):
    return F(np.float64(a) * np.float64(b) + np.float64(c))


def _sincos(x):
    x = F(x)
    k = F(np.rint(F(x * F(0.636619772367581343))))
    r = _fma(-k, F(1.5703125), x)
    r = _fma(-k, F(4.837512969970703125e-4), r)
    r = _fma(-k, F(7.54978995489188216e-8), r)
    z = F(r * r)
    sr = _fma(F(_fma(_fma(F(-1.9515295891e-4), z, F(8.3321608736e-3)), z, F(-1.6666654611e-1)) * z), r, r)
    cr = _fma(_fma(_fma(F(2.443315711809948e-5), z, F(-1.388731625493765e-3)), z, F(4.166664568298827e-2)),
              F(z * z), _fma(F(-0.5), z, F(1.0)))
    q = int(k)
    s1, c1 = (cr, sr) if q & 1 else (sr, cr)
    s = -s1 if q & 2 else s1
    c = -c1 if (q + 1) & 2 else c1
    return s, c


def test_half_angle_sincos_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(-7, 7, 4001), rng.uniform(-4096, 4096, 4000),
                         np.arange(-40, 41) * (np.pi / 2), [0.0, 1e-30, -1e-30, 4096.0, -4096.0]]).astype(np.float32)
    err_s = err_c = err_1 = 0.0
    for x in xs:
        s, c = _sincos(x)
        xd = np.float64(x)
        err_s = max(err_s, abs(np.float64(s) - np.sin(xd)))
        err_c = max(err_c, abs(np.float64(c) - np.cos(xd)))
        err_1 = max(err_1, abs(np.float64(s) ** 2 + np.float64(c) ** 2 - 1.0))
    # a few float32 ulps of 1; the reduction of |x| near 4096 adds the rounding of x / (pi / 2)
    assert err_s < 6e-7 and err_c < 6e-7 and err_1 < 1e-6, (err_s, err_c, err_1)


def test_half_angle_sincos_small_arguments_exact_order():
    # near zero the polynomials return x and 1 to the last bit, as the joint half angles of
    # converged problems need
    for x in (np.float32(1e-8), np.float32(-3e-5), np.float32(2.5e-3)):
        s, c = _sincos(x)
        assert abs(np.float64(s) - np.sin(np.float64(x))) <= 1.2e-7 * abs(np.float64(x)) + 1e-12
        assert abs(np.float64(c) - np.cos(np.float64(x))) <= 6e-8
