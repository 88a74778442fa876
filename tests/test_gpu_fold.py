"""The split-BF16 MFMA response fold against an FP64 matrix product, and the
error paths of a live context."""
import ctypes

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures, hip
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context, PpmError
from oracle_binding import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, nsx_path=fixtures.nsx_path(data_root, small=True), responses=("std", "std128"),
                nsx_config=small_nsx_cfg)
    yield c
    c.close()


@pytest.mark.parametrize("n_batch,n_phase", [(1, 32), (3, 17), (2, 128)])
def test_fold_matches_fp64(ctx, data_root, n_batch, n_phase):
    n_ch, e = ctx.response_info(0)
    rsp = np.loadtxt(data_root + "/tables/ism_inst/std/rsp.txt")
    rng = np.random.default_rng(n_phase)
    flux = rng.lognormal(-20, 2.0, (n_batch, n_phase, e.size))  # positive, wide dynamic range
    scale = rng.uniform(0.5, 2.0, n_batch) * 1e6
    d_flux = hip.DeviceArray.from_numpy(flux)
    d_cnt = hip.DeviceArray((n_batch, n_ch, n_phase))
    ctx.fold_device(0, n_batch, n_phase, d_flux.ptr, scale, d_cnt.ptr)
    got = d_cnt.numpy()
    ref = np.einsum("ce,bpe->bcp", rsp, flux) * scale[:, None, None]
    # hi/lo BF16 split of both operands keeps ~16 mantissa bits
    assert rel_err(got, ref) < 5e-5


def test_errors_on_live_context(ctx, data_root):
    st = S.cfg0_star()
    with pytest.raises(PpmError) as e:
        ctx.eval(st, abi.make_grid(0, 8, energies=[1.0]))
    assert e.value.status == abi.ERR_ARG
    with pytest.raises(PpmError) as e:  # flux energies differ from the response's E_obs
        ctx.eval(st, abi.make_grid(16, 8, energies=[1.0, 2.0], response=0))
    assert e.value.status == abi.ERR_ARG
    with pytest.raises(PpmError) as e:
        ctx.eval(st, abi.make_grid(16, 8, energies=[1.0], response=7))
    assert e.value.status == abi.ERR_ARG
    bad = S.cfg0_star()
    bad.mass_msun, bad.radius_eq_km = 3.0, 8.0    # R < R_s
    with pytest.raises(PpmError) as e:
        ctx.eval(bad, abi.make_grid(16, 8, energies=[1.0]))
    assert e.value.status == abi.ERR_ARG
    # NSX requested on a context without a grid
    with Context(data_root, nsx_path="", responses=()) as c2:
        with pytest.raises(PpmError) as e:
            c2.eval(S.cfg1_star(), abi.make_grid(16, 8, energies=[1.0]))
        assert e.value.status == abi.ERR_UNSUPPORTED
        # the context stays usable after an error
        f, _, _ = c2.eval(S.cfg0_star(), abi.make_grid(16, 8, energies=[1.0]))
        assert f.sum() > 0


def test_eval_device_outputs(ctx):
    g = abi.make_grid(32, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=0)
    st = S.cfg1_star()
    n_ch, e = ctx.response_info(0)
    d_flux = hip.DeviceArray((1, 16, e.size))
    d_cnt = hip.DeviceArray((1, n_ch, 16))
    stats = ctx.eval_device(st, g, d_flux.ptr, d_cnt.ptr)
    flux, counts, _ = ctx.eval(st, g)
    assert np.array_equal(d_flux.numpy(), flux) and np.array_equal(d_cnt.numpy(), counts)
    assert stats.kernel_ms > 0 and stats.n_eval_ref > 0
