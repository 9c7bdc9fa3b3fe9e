"""The C ABI (include/cfm.h) without a GPU: the library loads, exports every
declared symbol, validates arguments, and its host-side generators agree with
the oracle bit for bit."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import REPO


def declared_symbols():
    src = open(os.path.join(REPO, "include", "cfm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cfm_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported(cfm):
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(cfm.lib, name), name
    assert set(names) == set(cfm.EXPORTS)


def test_version_and_status_strings(cfm):
    assert "gfx950" in cfm.version()
    assert cfm.status_str(cfm.OK) == "ok"
    assert "converged" in cfm.status_str(cfm.ENOTCONV)
    assert cfm.status_str(99) == "unknown status"


@pytest.mark.parametrize("dims", [(4, 4, 4), (7, 3, 5), (128, 2, 1)])
def test_poisson_generator_matches_oracle(cfm, orc, dims):
    rp, col, val = cfm.poisson3d(*dims)
    A = orc.poisson3d(*dims)
    assert np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col) and np.array_equal(val, A.val)


def test_uniform_generator_matches_oracle(cfm, orc):
    for seed, sid in ((0, 0), (12345, 7), (2**63 + 5, 255)):
        assert np.array_equal(cfm.uniform(seed, sid, 4096), orc.uniform(seed, sid, 4096))


def test_init_fields_match_oracle(cfm, orc):
    assert np.array_equal(cfm.init_field(19, 13), orc.init2d(19, 13))
    assert np.array_equal(cfm.init_field(9, 8, 7), orc.init3d(9, 8, 7))
    assert np.array_equal(cfm.init_field(9, 8, 7, cfm.INIT_REF_QUIRK), orc.init3d(9, 8, 7, orc.INIT_REF_QUIRK))
    f32 = cfm.init_field(9, 8, 7, prec=cfm.F32)
    assert f32.dtype == np.float32 and np.array_equal(f32, orc.init3d(9, 8, 7, dtype=np.float32))
    assert np.array_equal(cfm.init_field(19, 13, prec=cfm.F32), orc.init2d(19, 13, dtype=np.float32))
    assert np.array_equal(cfm.init_field(9, 8, 7, cfm.INIT_REF_QUIRK, prec=cfm.F32),
                          orc.init3d(9, 8, 7, orc.INIT_REF_QUIRK, dtype=np.float32))


def test_f32_init_uses_the_float_rounded_pi(cfm, orc):
    # The reference's PI is `real` (float): static_cast<real>(2*asin(1)), widened to double
    # inside sin(2.0*pi*iy/(ny-1)) (Stencil/jacobi2D/include/common.h:24, src/common.cu:20).
    # Expected values from CPython's math.sin (the C library's sin), independent of both.
    ny = 1024
    pi_f = float(np.float32(2.0 * math.asin(1.0)))
    want = np.array([math.sin(2.0 * pi_f * iy / (ny - 1)) for iy in range(ny)]).astype(np.float32)
    got = cfm.init_field(8, ny, prec=cfm.F32)
    assert np.array_equal(got[:, 0], want) and np.array_equal(got[:, -1], want)
    assert np.array_equal(orc.init2d(8, ny, dtype=np.float32)[:, 0], want)
    # the double-precision pi gives different float values on most rows: the test can tell them apart
    f64pi = cfm.init_field(8, ny)[:, 0].astype(np.float32)
    assert np.count_nonzero(f64pi != want) > ny // 4
    # 3D faces: sin(2 pi iy/(ny-1)) * sin(2 pi iz/(nz-1)) in double with the same pi, then float
    nz, ny3 = 11, 9
    want3 = np.array([[math.sin(2.0 * pi_f * iy / (ny3 - 1)) * math.sin(2.0 * pi_f * iz / (nz - 1))
                       for iy in range(ny3)] for iz in range(nz)]).astype(np.float32)
    g3 = cfm.init_field(6, ny3, nz, prec=cfm.F32)
    assert np.array_equal(g3[:, :, 0], want3) and np.array_equal(g3[:, :, -1], want3)


def test_argument_validation(cfm):
    bad = [dict(nx=2, ny=5), dict(nx=5, ny=5, nz=2), dict(nx=5, ny=5, niter=-1), dict(nx=5, ny=5, prec=7),
           dict(nx=5, ny=5, bc=3), dict(nx=5, ny=5, variant=9)]
    for kw in bad:
        p = cfm.jacobi_params(**kw)
        h = ctypes.c_void_p()
        assert cfm.lib.cfm_jacobi_plan_create(ctypes.byref(p), ctypes.byref(h)) == cfm.EINVAL, kw
        assert not h.value
    assert cfm.lib.cfm_jacobi_plan_create(None, None) == cfm.EINVAL
    assert cfm.lib.cfm_jacobi_init_field(2, 5, 1, 0, 0, None) == cfm.EINVAL
    rp = np.array([0, 1], dtype=np.int32)
    col = np.array([0], dtype=np.int32)
    val = np.array([1.0])
    p = cfm.cg_params(method=9)
    h = ctypes.c_void_p()
    assert cfm.lib.cfm_cg_plan_create(1, 1, rp.ctypes.data, col.ctypes.data, val.ctypes.data, 0, 0,
                                      ctypes.byref(p), ctypes.byref(h)) == cfm.EINVAL
    p = cfm.cg_params(method=cfm.CG_IR_F32, mode=cfm.CG_DISCRETE)
    assert cfm.lib.cfm_cg_plan_create(1, 1, rp.ctypes.data, col.ctypes.data, val.ctypes.data, 0, 0,
                                      ctypes.byref(p), ctypes.byref(h)) == cfm.ENOTSUP


def test_no_device_is_an_error_not_a_fallback(cfm):
    if cfm.device_count() > 0:
        pytest.skip("a GPU is visible")
    p = cfm.jacobi_params(8, 8, 8, niter=1)
    h = ctypes.c_void_p()
    assert cfm.lib.cfm_jacobi_plan_create(ctypes.byref(p), ctypes.byref(h)) == cfm.EDEVICE
    with pytest.raises(cfm.CfmError):
        cfm.jacobi(np.zeros((8, 8, 8)), 1)


def test_varcoef_generator_matches_oracle(cfm, orc):
    for sid in range(12):
        A, f, dims = orc.varcoef3d(42, sid, 4, 11)
        rp, col, val, f2, d2 = cfm.varcoef3d(42, sid, 4, 11)
        assert d2 == dims
        assert np.array_equal(rp, A.rowptr) and np.array_equal(col, A.col)
        assert np.array_equal(val, A.val) and np.array_equal(f2, f)


def test_block_diag_layout(cfm):
    s1 = (np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), np.array([2.0, 3.0]))
    s2 = (np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32), np.array([4.0, -1, -1, 4.0]))
    srp, rp, col, val = cfm.block_diag([s1, s2])
    assert srp.tolist() == [0, 2, 4]
    assert rp.tolist() == [0, 1, 2, 4, 6]
    assert col.tolist() == [0, 1, 2, 3, 2, 3]


def test_unet_host_side(cfm):
    assert cfm.unet_param_count() == sum(co * ci * 27 + co for ci, co in cfm.UNET_CONVS) + 17
    p = cfm.unet_init_params(5)
    assert np.array_equal(p, cfm.unet_init_params(5))
    assert not np.array_equal(p, cfm.unet_init_params(6))
    off = 0
    for ci, co in cfm.UNET_CONVS:
        a = np.sqrt(6.0 / (27 * ci))
        w = p[off: off + co * ci * 27]
        assert np.all(np.abs(w) <= a) and np.abs(w).max() > 0.9 * a
        assert abs(float(w.mean())) < 0.1 * a
        off += co * ci * 27
        assert np.all(p[off: off + co] == 0)
        off += co
    assert np.all(np.abs(p[off: off + 16]) <= 0.1 * np.sqrt(6.0 / 16)) and p[off + 16] == 0


@pytest.mark.parametrize("n,batch", [(6, 1), (0, 1), (8, 0), (2048, 64)])
def test_unet_create_rejects_bad_shapes(cfm, n, batch):
    with pytest.raises(cfm.CfmError) as e:
        cfm.UNet(n, batch)
    assert e.value.status == cfm.EINVAL


def test_varcoef_coefficients_match_matrix(cfm):
    """The per-cell k reproduces every matrix entry: off-diagonal
    -2 k_i k_j / (k_i + k_j), boundary faces k_i, diagonal the sum of the
    six faces in the order z-, y-, x-, x+, y+, z+."""
    rp, col, val, f, dims = cfm.varcoef3d(11, 4, 3, 7)
    k = cfm.varcoef3d_coef(11, 4, 3, 7)
    assert k.shape == (dims[2], dims[1], dims[0])
    nx, ny, nz = dims
    kf = k.reshape(-1)
    for i in range(kf.size):
        ix, iy, iz = i % nx, (i // nx) % ny, i // (nx * ny)
        nbs = [i - nx * ny if iz > 0 else -1, i - nx if iy > 0 else -1, i - 1 if ix > 0 else -1,
               i + 1 if ix < nx - 1 else -1, i + nx if iy < ny - 1 else -1, i + nx * ny if iz < nz - 1 else -1]
        cf = [(2.0 * kf[i] * kf[j]) / (kf[i] + kf[j]) if j >= 0 else kf[i] for j in nbs]
        diag = 0.0
        for c in cf:
            diag = diag + c
        row = dict(zip(col[rp[i]:rp[i + 1]], val[rp[i]:rp[i + 1]]))
        assert row.pop(i) == diag
        for j, c in zip(nbs, cf):
            if j >= 0:
                assert row.pop(j) == -c
        assert not row
    lk = np.log(cfm.varcoef3d_coef(11, 5, 32, 32))
    assert abs(lk.mean()) < 0.01 and abs(lk.std() - 1.0) < 0.01


def test_refuses_foreign_hip_runtime():
    """Importing cfm after a library that brought its own HIP runtime (the
    torch wheel) fails loudly instead of running kernels on that runtime."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import torch\n"
            "try:\n    import cfm\nexcept ImportError as e:\n    print('refused', 'torch' in str(e))\n"
            "else:\n    print('loaded')\n") % os.path.join(REPO, "cpu-free-model_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    torch_dir = subprocess.run([sys.executable, "-c", "import torch, os; print(os.path.dirname(torch.__file__))"],
                               capture_output=True, text=True, timeout=300).stdout.strip()
    if not os.path.exists(os.path.join(torch_dir, "lib", "libamdhip64.so")):
        pytest.skip("this torch build does not bundle a HIP runtime")
    assert out.stdout.strip() == "refused True", out.stdout + out.stderr
