"""The blackbody closed form of the NSX path on the device (see
test_blackbody_pin.py): with the Planck-filled NSX grid, the device's NSX
path, plain and batched, equals its blackbody path in every energy column
within the energy axis' interpolation error of the Planck function (plus
2e-6 for the FP32 evaluation), and the FP64 oracle within 1e-4."""
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context
from oracle_binding import Oracle, rel_err
from test_blackbody_pin import check_columns, column_bounds, grid, star_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bb_root(data_root):
    fixtures.ensure_nsx_bb(data_root)
    return data_root


@pytest.mark.parametrize("method", [abi.METHOD_PLAIN, abi.METHOD_BATCHED])
@pytest.mark.parametrize("mode", [abi.NSX_2L2C_LE, abi.NSX_4L])
def test_device_nsx_path_reproduces_blackbody(bb_root, small_nsx_cfg, method, mode):
    bb, nsx = star_pair(mode)
    g = grid()
    g.method = method
    print("case: blackbody-filled NSX grid, interpolant %d, method %d" % (mode, method), flush=True)
    with Context(bb_root, nsx_path=fixtures.nsx_bb_path(bb_root), nsx_config=small_nsx_cfg) as ctx:
        f_bb, _, s_bb = ctx.eval([bb], g)
        f, _, s = ctx.eval([nsx], g)
    assert s.n_active == s_bb.n_active > 0
    en = S.geomspace(0.2, 12.0, 64)
    bound = column_bounds(en, bb.spot[0].kT_keV, cubic=mode != abi.NSX_4L)
    r = check_columns(f[0], f_bb[0], bound, 2e-6)
    print("  NSX vs blackbody: %.2e at 0.2 keV, %.2e at 1 keV" % (r[0], r[en <= 1.0].max()), flush=True)
    orc = Oracle(bb_root, nsx_path=fixtures.nsx_bb_path(bb_root), nsx_cfg=small_nsx_cfg)
    try:
        ref, _ = orc.profile(nsx, g, threads=16)
    finally:
        orc.close()
    assert rel_err(f[0], ref) <= 1e-4
