"""cfm_jacobi_run_noalloc (the entry point that allocates nothing) against the
plan entry point cfm_jacobi_run on the same fields: field and update norms bit
for bit, two calls in a row with the second one smaller (the device scratch
array keeps the first call's contents), plus its host-side refusals."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=shape).astype(dtype)


def _noalloc(gpu, init, niter, bc, norm_every, **tuning):
    prec = gpu.F64 if init.dtype == np.float64 else gpu.F32
    nz, ny, nx = (1,) + init.shape if init.ndim == 2 else init.shape
    p = gpu.jacobi_params(nx, ny, nz, niter, norm_every, prec, bc, **tuning)
    with gpu.DeviceBuffer.of(init) as field, gpu.DeviceBuffer(init.nbytes) as work:
        norms, t = gpu.jacobi_noalloc(p, field, work)
        out = field.download(init.dtype, init.shape)
    return out, norms, t


# (shape, dtype, niter, bc, norm_every, tuning): the first case of each pair is
# the larger one; the runs cover one, two and four sweeps per pass, odd and
# even pass counts (the result copied back from the work array or not),
# periodic halos and one launch per sweep
PAIRS = [
    (((34, 40, 66), np.float64, 10, 1, 5, {}), ((9, 13, 20), np.float64, 7, 1, 3, {})),
    (((66, 130), np.float64, 9, 1, 2, {}), ((18, 11), np.float64, 5, 1, 1, {})),
    (((20, 24, 70), np.float64, 6, 0, 2, {}), ((10, 9, 7), np.float64, 3, 0, 1, {})),
    (((40, 37, 100), np.float32, 11, 1, 4, {}), ((12, 70, 33), np.float32, 4, 1, 0, {})),
    (((130, 70), np.float64, 5, 0, 1, {"variant": 0}), ((30, 20), np.float64, 3, 0, 1, {"variant": 0})),
]


@pytest.mark.parametrize("pair", range(len(PAIRS)))
def test_noalloc_matches_plan_two_calls(gpu, pair):
    for k, (shape, dtype, niter, bc, norm_every, tuning) in enumerate(PAIRS[pair]):
        tuning = dict(tuning)
        init = _rand(shape, dtype, 100 * pair + k)
        variant = tuning.pop("variant", gpu.PERSISTENT)
        ref, rnorms, rt = gpu.jacobi(init, niter, bc=bc, norm_every=norm_every, variant=variant, **tuning)
        out, norms, t = _noalloc(gpu, init, niter, bc, norm_every, variant=variant, **tuning)
        assert t["outer_iterations"] == rt["outer_iterations"]
        assert np.array_equal(out, ref), (pair, k)
        assert np.array_equal(norms, rnorms), (pair, k)


def test_noalloc_refuses_what_the_array_cannot_hold(gpu):
    init = _rand((10, 9, 7), np.float64, 5)
    with pytest.raises(gpu.CfmError) as e:
        _noalloc(gpu, init, 400000, 1, 1)  # 400000 norm slots x every workgroup > the scratch array
    assert e.value.status == gpu.EINVAL and "scratch" in str(e.value)
    with pytest.raises(gpu.CfmError) as e:
        _noalloc(gpu, init, 4, 1, 1, phase_timers=1)
    assert e.value.status == gpu.EINVAL
    # the array still serves a normal call afterwards
    ref, rnorms, _ = gpu.jacobi(init, 4, bc=1, norm_every=1)
    out, norms, _ = _noalloc(gpu, init, 4, 1, 1)
    assert np.array_equal(out, ref) and np.array_equal(norms, rnorms)
