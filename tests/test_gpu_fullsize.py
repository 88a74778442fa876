"""Parity at the full sizes of the workloads, where the FP64 oracle would take
hours per profile, through properties that hold at any size:

* ring + crescent star at N_side 2048, 128 bins x 2 GL nodes, 360 energies
  with the 300x360 fold: the device's active-patch count equals the oracle's
  (the survey's 3,259,051); the batched path equals the plain one (two
  independent accumulation schemes: Kahan-compensated FP32 per bin against
  phase-shift dedupe with FP64 prefix sums); the ring equals the outer cap
  minus the inner cap (the surface integral is a sum over pixels); the counts
  equal the oracle's FP64 fold of the device flux;
* the single-cap star at N_side 1024, 64 bins x 2 GL nodes, 360 energies
  (the bench grid): active count, batched == plain, fold;
* a spot on the spin axis seen pole-on gives a phase-independent profile.

Tolerance: 1e-4 max relative error per bin above 1e-6 of the peak, as for the
oracle comparisons (1e-5 between the two device paths)."""
import copy

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context
from oracle_binding import Oracle, count_active, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-4
TOL_PATHS = 1e-5


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                responses=("std", "std128"), nsx_config=small_nsx_cfg)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc(data_root, small_nsx_cfg):
    o = Oracle(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
               responses=("std", "std128"), nsx_cfg=small_nsx_cfg)
    yield o
    o.close()


def grid(n_side, n_phase, method):
    # energies: the 360 E_obs of response 0 (300 channels x 360)
    return abi.make_grid(n_side, n_phase, sampling=abi.SAMPLE_GL, n_gl=2, response=0, method=method)


def one_spot(star, spot):
    st = copy.copy(star)
    st.n_spots = 1
    st.spot[0] = spot
    return st


@pytest.mark.parametrize("cfg", ["ring-crescent N2048", "cap N1024"])
def test_full_size_paths_agree_and_fold(ctx, orc, cfg):
    star, n_side, n_phase, n_act = ((S.cfg2_star(), 2048, 128, 3259051) if cfg.startswith("ring")
                                    else (S.cfg1_star(), 1024, 64, 195598))
    gp, gb = grid(n_side, n_phase, abi.METHOD_PLAIN), grid(n_side, n_phase, abi.METHOD_BATCHED)
    assert count_active(star, gp) == n_act
    print("case: active count,", cfg, flush=True)
    assert ctx.count_active(star, gp) == n_act
    print("case: plain,", cfg, flush=True)
    fp, cp, sp = ctx.eval([star], gp)
    print("case: batched,", cfg, flush=True)
    fb, cb, sb = ctx.eval([star], gb)
    assert sp.n_active == sb.n_active == n_act
    assert np.abs(fp).max() > 0
    e_flux, e_counts = rel_err(fb[0], fp[0]), rel_err(cb[0], cp[0])
    print("batched vs plain: flux %.2e counts %.2e" % (e_flux, e_counts), flush=True)
    assert e_flux <= TOL_PATHS and e_counts <= TOL_PATHS
    # the fold of the device flux, in FP64 by the oracle, against the MFMA fold
    for f, c in ((fp[0], cp[0]), (fb[0], cb[0])):
        e = rel_err(c, orc.fold(f, 0, star.exposure_s / n_phase))
        print("fold vs FP64 fold: %.2e" % e, flush=True)
        assert e <= TOL


def test_full_size_ring_is_outer_minus_inner_cap(ctx):
    """N_side 2048: the ring of the ring + crescent star (0.15-0.4 rad at
    35 deg) against the two caps it is cut from, evaluated separately."""
    star = S.cfg2_star()
    ring = star.spot[0]
    outer = abi.make_spot(ring.theta1, ring.angrad1, ring.kT_keV)
    inner = abi.make_spot(ring.theta2, ring.angrad2, ring.kT_keV)
    g = grid(2048, 128, abi.METHOD_PLAIN)
    out = {}
    for name, sp in (("ring", ring), ("outer", outer), ("inner", inner)):
        print("case: N 2048 single spot,", name, flush=True)
        out[name], _, stats = ctx.eval([one_spot(star, sp)], g, counts=False)
        assert stats.n_active > 0
    e = rel_err(out["ring"][0], out["outer"][0] - out["inner"][0])
    print("ring vs outer - inner: %.2e" % e, flush=True)
    assert e <= TOL


def test_full_size_pole_on_is_phase_independent(ctx):
    """N_side 1024, 64 bins: a cap centred on the spin axis seen from the spin
    axis looks the same at every rotational phase."""
    st = S.cfg1_star()
    st.inclination_rad = 0.0
    st.spot[0] = abi.make_spot(0.0, 0.3, st.spot[0].kT_keV)
    for method in (abi.METHOD_PLAIN, abi.METHOD_BATCHED):
        print("case: pole-on N 1024, method", method, flush=True)
        f, _, stats = ctx.eval([st], grid(1024, 64, method), counts=False)
        assert stats.n_active > 0
        e = rel_err(f[0], np.broadcast_to(f[0].mean(axis=0), f[0].shape))
        print("largest deviation from the phase mean: %.2e" % e, flush=True)
        assert e <= TOL_PATHS
