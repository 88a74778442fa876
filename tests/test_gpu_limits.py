"""The ABI's largest grids (ppm.h: N_side <= 8192, n_phase <= 4096,
n_fine <= 8192).  At the finest phase resolution (4096 bins, 8192 fine
phases) the batched path matches the FP64 oracle and the plain path, whose
ring tables would not fit one workgroup's LDS, refuses the grid with
PPM_ERR_ARG instead of launching.  At N_side 8192 the two paths agree on a
small spot, and the active-pixel count gives the spot's solid angle."""
import math

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd.lib import Context, PpmError
from oracle_binding import Oracle, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, nsx_path=fixtures.nsx_path(data_root, small=True), responses=("std",),
                nsx_config=small_nsx_cfg)
    yield c
    c.close()


def bb_star(theta, rho):
    return abi.make_star([abi.make_spot(theta, rho, 0.35)], mass_msun=1.4, radius_eq_km=12.0, spin_hz=400.0,
                         inclination_rad=1.0, distance_km=abi.KPC_KM, atmosphere=abi.ATM_BLACKBODY)


def test_finest_phase_grid(ctx, data_root, small_nsx_cfg):
    st = bb_star(1.0, 0.8)
    e = [0.5, 3.0]
    gb = abi.make_grid(16, 4096, n_fine=8192, sampling=abi.SAMPLE_GL, n_gl=2, energies=e, method=abi.METHOD_BATCHED)
    print("case: N 16, 4096 bins x 2 GL, n_fine 8192, batched", flush=True)
    f, _, s = ctx.eval(st, gb)
    orc = Oracle(data_root, nsx_path=fixtures.nsx_path(data_root, small=True), responses=("std",),
                 nsx_cfg=small_nsx_cfg)
    ref, info = orc.profile(st, gb, threads=16)
    orc.close()
    assert s.n_active == info.n_active > 0
    assert rel_err(f[0], ref) <= TOL, rel_err(f[0], ref)
    gp = abi.make_grid(16, 4096, n_fine=8192, sampling=abi.SAMPLE_GL, n_gl=2, energies=e, method=abi.METHOD_PLAIN)
    print("case: the same grid on the plain path (refused)", flush=True)
    with pytest.raises(PpmError) as err:
        ctx.eval(st, gp)
    assert err.value.status == abi.ERR_ARG


def test_finest_healpix_grid(ctx):
    rho = 0.01
    st = bb_star(1.0, rho)
    e = [0.3, 1.0, 2.0, 6.0]
    out = {}
    for m in (abi.METHOD_PLAIN, abi.METHOD_BATCHED):
        g = abi.make_grid(8192, 16, sampling=abi.SAMPLE_GL, n_gl=2, energies=e, method=m)
        print("case: N 8192, spot 0.01 rad, method %d" % m, flush=True)
        out[m] = ctx.eval(st, g)
    (fp, _, sp), (fb, _, sb) = out[abi.METHOD_PLAIN], out[abi.METHOD_BATCHED]
    assert sp.n_active == sb.n_active > 0
    assert np.abs(fp).max() > 0 and rel_err(fb[0], fp[0]) <= 1e-5
    # HEALPix pixels have equal area pi / (3 N^2): the count measures the cap
    cap = 2 * math.pi * (1 - math.cos(rho))
    assert abs(sp.n_active * math.pi / (3 * 8192.0 ** 2) / cap - 1) < 0.01
