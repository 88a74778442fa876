"""ppm_eval_device_noalloc, the entry point that allocates no device memory
(its buffers live in the library's fixed per-GPU scratch array), against
ppm_eval_device on the plain path: two calls in a row, the second smaller,
bit-identical outputs (the same kernels in the same order) and both within
1e-4 of the FP64 oracle; a profile larger than the scratch is refused before
anything is launched."""
import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures, hip
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context, PpmError
from oracle_binding import Oracle, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, nsx_path=fixtures.nsx_path(data_root, small=True), responses=("std", "std128"),
                nsx_config=small_nsx_cfg)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc(data_root, small_nsx_cfg):
    o = Oracle(data_root, nsx_path=fixtures.nsx_path(data_root, small=True), responses=("std", "std128"),
               nsx_cfg=small_nsx_cfg)
    yield o
    o.close()


def _run(ctx, star, g, noalloc):
    n_ch = ctx.response_info(g.response)[0]
    ne = ctx.n_energy(g)
    d_f = hip.DeviceArray((g.n_phase, ne))
    d_c = hip.DeviceArray((n_ch, g.n_phase))
    if noalloc:
        st = ctx.eval_device_noalloc(star, g, d_f.ptr, d_c.ptr)
    else:
        st = ctx.eval_device([star], g, d_f.ptr, d_c.ptr)
    hip.synchronize()
    return d_f.numpy(), d_c.numpy(), st


def test_noalloc_equals_eval_device_twice(ctx, orc):
    cases = [("cap N256 64x2 GL, 360 E", S.cfg1_star(),
              abi.make_grid(256, 64, sampling=abi.SAMPLE_GL, n_gl=2, response=0, method=abi.METHOD_PLAIN)),
             ("cap N64 16x2 GL, 128 E (smaller)", S.cfg0_star(),
              abi.make_grid(64, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=1, method=abi.METHOD_PLAIN))]
    for name, star, g in cases:
        print("case: noalloc vs eval_device,", name, flush=True)
        f0, c0, s0 = _run(ctx, star, g, noalloc=False)
        f1, c1, s1 = _run(ctx, star, g, noalloc=True)
        assert s1.n_active == s0.n_active > 0
        assert np.array_equal(f1, f0) and np.array_equal(c1, c0), name
        ref, info = orc.profile(star, g, threads=8)
        cref = orc.fold(ref, g.response, star.exposure_s / g.n_phase)
        ef, ec = rel_err(f1, ref), rel_err(c1, cref)
        print("  vs oracle: flux %.2e counts %.2e" % (ef, ec), flush=True)
        assert info.n_active == s1.n_active
        assert ef <= TOL and ec <= TOL


def test_noalloc_refuses_beyond_scratch(ctx):
    star = S.cfg2_star()
    g = abi.make_grid(2048, 128, sampling=abi.SAMPLE_GL, n_gl=2, response=0, method=abi.METHOD_PLAIN)
    d_f = hip.DeviceArray((g.n_phase, 360))
    with pytest.raises(PpmError) as ei:
        ctx.eval_device_noalloc(star, g, d_f.ptr, None)
    assert ei.value.status == abi.ERR_NOMEM
    assert "scratch" in str(ei.value)
    g.method = abi.METHOD_BATCHED
    with pytest.raises(PpmError) as ei:
        ctx.eval_device_noalloc(star, g, d_f.ptr, None)
    assert ei.value.status == abi.ERR_ARG
