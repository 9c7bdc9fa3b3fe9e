"""Pins of the Jacobi CPU oracle (oracle/jacobi_oracle.c).

The reference ships no golden vectors or known-answer tests for this path, so
the oracle is pinned by what the mathematics fixes, chosen so that a dropped
term, a wrong neighbour index, a transposed axis, a changed summation order,
a wrong halo copy or a wrong swap each fail at least one check:
  * niter = 0 is the identity; a constant field is a fixed point (bitwise);
  * a unit impulse spreads exactly 1/4 (2D) or fl(1/6) (3D) to each neighbour,
    on non-square grids, and gives a closed-form update norm;
  * crafted neighbour values (1e16, 1, -1e16, 1) make the result depend on the
    left-to-right order x+1, x-1, y+1, y-1 (, z+1, z-1) of the reference
    expression (Stencil/jacobi2D/src/common.cu:37, Stencil/jacobi3D/src/common.cu:43);
  * periodic mode reproduces the reference's row/plane copies
    (Stencil/jacobi2D/src/common.cu:234-242);
  * runs compose (niter = a + b), which pins the swap bookkeeping;
  * with Dirichlet data sin(theta*y) (2D) or sin*sin (3D) the iteration
    converges to the exact discrete harmonic function
    sin(theta y) cosh(mu (x - c)) / cosh(mu c), 2 cosh(mu) - 2 = sum(2 - 2 cos theta);
  * mirror symmetry in x and the discrete maximum principle.
"""
import math

import numpy as np
import pytest


def _run(orc, a, niter, bc=1, norm_every=0):
    return orc.jacobi(a, niter, bc, norm_every)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(11, 18), (7, 9, 10)])
def test_niter_zero_identity(orc, dtype, shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape).astype(dtype)
    for bc in (0, 1):
        out, _ = _run(orc, a, 0, bc)
        assert np.array_equal(out, a)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(11, 18), (7, 9, 10)])
def test_constant_field_fixed_point(orc, dtype, shape):
    a = np.ones(shape, dtype=dtype)
    for bc in (0, 1):
        out, norms = _run(orc, a, 5, bc, norm_every=1)
        assert np.array_equal(out, a)
        assert np.all(norms == 0.0)


def test_impulse_2d(orc):
    ny, nx = 9, 13
    a = np.zeros((ny, nx))
    a[4, 6] = 1.0
    out, norms = _run(orc, a, 1, 1, norm_every=1)
    exp = np.zeros_like(a)
    exp[4, 7] = exp[4, 5] = exp[5, 6] = exp[3, 6] = 0.25
    assert np.array_equal(out, exp)
    assert norms[0] == math.sqrt(1.0 + 4 * 0.0625)


def test_impulse_3d(orc):
    nz, ny, nx = 7, 9, 11
    a = np.zeros((nz, ny, nx))
    a[3, 4, 5] = 1.0
    out, norms = _run(orc, a, 1, 1, norm_every=1)
    c = 1.0 / 6.0
    exp = np.zeros_like(a)
    for dz, dy, dx in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)):
        exp[3 + dz, 4 + dy, 5 + dx] = c
    assert np.array_equal(out, exp)
    # the sum order over rows is the oracle's own; the value is fixed to rounding
    assert norms[0] == pytest.approx(math.sqrt(math.fsum([1.0] + [c * c] * 6)), rel=4e-16)


def _left_to_right(vals, coeff):
    s = vals[0]
    for v in vals[1:]:
        s = s + v
    return coeff * s


def test_summation_order_2d(orc):
    # E, W, N, S = a[iy][ix+1], a[iy][ix-1], a[iy+1][ix], a[iy-1][ix]
    E, W, N, S = 1e16, 1.0, -1e16, 1.0
    a = np.zeros((5, 5))
    a[2, 3], a[2, 1], a[3, 2], a[1, 2] = E, W, N, S
    out, _ = _run(orc, a, 1, 1)
    want = _left_to_right([E, W, N, S], 0.25)
    assert out[2, 2] == want == 0.25
    # the check discriminates: other orders give other values
    assert want != _left_to_right([E, N, W, S], 0.25)
    assert want != _left_to_right([S, N, W, E], 0.25)


def test_summation_order_3d(orc):
    # order x+1, x-1, y+1, y-1, z+1, z-1
    vals = [1e16, 1.0, 1.0, -1e16, 1.0, 1.0]
    a = np.zeros((5, 5, 5))
    idx = [(2, 2, 3), (2, 2, 1), (2, 3, 2), (2, 1, 2), (3, 2, 2), (1, 2, 2)]
    for v, (z, y, x) in zip(vals, idx):
        a[z, y, x] = v
    out, _ = _run(orc, a, 1, 1)
    c = 1.0 / 6.0
    want = _left_to_right(vals, c)
    assert out[2, 2, 2] == want
    assert want != _left_to_right(vals[::-1], c)
    assert want != _left_to_right([vals[4], vals[5], vals[0], vals[1], vals[2], vals[3]], c)


def test_summation_order_f32(orc):
    E, W, N, S = np.float32(1e8), np.float32(1), np.float32(-1e8), np.float32(1)
    a = np.zeros((5, 5), dtype=np.float32)
    a[2, 3], a[2, 1], a[3, 2], a[1, 2] = E, W, N, S
    out, _ = _run(orc, a, 1, 1)
    want = np.float32(0.25) * (((E + W) + N) + S)
    assert out[2, 2] == want
    assert want != np.float32(0.25) * ((E + N) + (W + S))


def test_periodic_copies_2d(orc):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((10, 12))
    for niter in (1, 2, 3):
        out, _ = _run(orc, a, niter, 0)
        assert np.array_equal(out[0], out[-2])
        assert np.array_equal(out[-1], out[1])
    # Dirichlet keeps the faces
    out, _ = _run(orc, a, 3, 1)
    assert np.array_equal(out[0], a[0]) and np.array_equal(out[-1], a[-1])
    assert np.array_equal(out[:, 0], a[:, 0]) and np.array_equal(out[:, -1], a[:, -1])


def test_periodic_copies_3d(orc):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, 7, 8))
    for niter in (1, 2):
        out, _ = _run(orc, a, niter, 0)
        assert np.array_equal(out[0], out[-2])
        assert np.array_equal(out[-1], out[1])


def test_periodic_wraps_neighbour(orc):
    # In periodic mode the first interior row sees the last interior row of the
    # previous sweep through row 0: an impulse on row ny-2 reaches row 1 after
    # two sweeps (copy, then stencil).
    ny, nx = 8, 7
    a = np.zeros((ny, nx))
    a[ny - 2, 3] = 1.0
    out1, _ = _run(orc, a, 1, 0)
    assert out1[0, 3] == out1[ny - 2, 3]
    out2, _ = _run(orc, a, 2, 0)
    assert out2[1, 2] == 0.0625 and out2[1, 4] == 0.0625
    outd, _ = _run(orc, a, 2, 1)
    assert np.all(outd[1] == 0.0)


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("shape", [(11, 18), (7, 9, 10)])
def test_runs_compose(orc, bc, shape):
    rng = np.random.default_rng(4)
    a = rng.standard_normal(shape)
    full, _ = _run(orc, a, 5, bc)
    part, _ = _run(orc, a, 2, bc)
    rest, _ = _run(orc, part, 3, bc)
    assert np.array_equal(full, rest)


def test_harmonic_limit_2d(orc):
    ny, nx = 17, 15
    a = orc.init2d(nx, ny)
    out, norms = _run(orc, a, 4000, 1, norm_every=1000)
    theta = 2 * math.pi / (ny - 1)
    mu = math.acosh(2.0 - math.cos(theta))
    cx = (nx - 1) / 2
    iy, ix = np.mgrid[0:ny, 0:nx]
    exact = np.sin(theta * iy) * np.cosh(mu * (ix - cx)) / math.cosh(mu * cx)
    assert np.max(np.abs(out - exact)) < 1e-10
    assert norms[0] == 1.0 and np.all(norms[1:] < 1e-12)
    # first sweep: columns 1 and nx-2 become 0.25 sin(theta y); sum of sin^2 over a
    # period of 16 is 8, so the norm is 0.25 * sqrt(2 * 8) = 1; later updates vanish.


def test_harmonic_limit_3d(orc):
    nz, ny, nx = 9, 11, 13
    a = orc.init3d(nx, ny, nz)
    out, _ = _run(orc, a, 3000, 1)
    ty, tz = 2 * math.pi / (ny - 1), 2 * math.pi / (nz - 1)
    mu = math.acosh(1.0 + (1.0 - math.cos(ty)) + (1.0 - math.cos(tz)))
    cx = (nx - 1) / 2
    iz, iy, ix = np.mgrid[0:nz, 0:ny, 0:nx]
    exact = np.sin(ty * iy) * np.sin(tz * iz) * np.cosh(mu * (ix - cx)) / math.cosh(mu * cx)
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.max(np.abs(out[inner] - exact[inner])) < 1e-10


def test_mirror_symmetry_and_max_principle(orc):
    a = orc.init2d(21, 13)
    out, _ = _run(orc, a, 50, 1)
    assert np.array_equal(out, out[:, ::-1])
    b = orc.init3d(12, 9, 8)
    out3, _ = _run(orc, b, 20, 1)
    assert np.array_equal(out3, out3[:, :, ::-1])
    lo, hi = a.min(), a.max()
    assert out[1:-1, 1:-1].min() >= lo and out[1:-1, 1:-1].max() <= hi


def test_init_fields(orc):
    a = orc.init2d(6, 5)
    assert np.array_equal(a[:, 0], np.sin(2.0 * math.pi * np.arange(5) / 4))
    assert np.array_equal(a[:, 0], a[:, -1])
    assert np.all(a[:, 1:-1] == 0)
    q = orc.init3d(5, 4, 3, orc.INIT_REF_QUIRK)
    iz, iy = np.mgrid[0:3, 0:4]
    assert np.array_equal(q[:, :, 0], (iz - iy - 4).astype(float))
    assert np.all(q[:, :, 1:] == 0)


def test_f32_tracks_f64(orc):
    a = orc.init3d(16, 12, 10)
    o64, n64 = _run(orc, a, 40, 1, norm_every=10)
    o32, n32 = _run(orc, a.astype(np.float32), 40, 1, norm_every=10)
    assert np.max(np.abs(o32.astype(np.float64) - o64)) < 1e-6
    assert np.allclose(n32, n64, rtol=1e-5)
