"""Closed form of the NSX path: an NSX-format grid filled with the isotropic
Planck intensity alone (logInuT3 = log10(2 k^3 x^3 / (h^2 c^2 (e^x - 1))),
x = E/kT, independent of logT, log g and mu; ppm_synth nsx_bb) makes the
NSX path of the oracle evaluate the blackbody path's intensity
(cpu/unit.hpp:28-31 against cpu/nsx.hpp:248-296 with
cpu/devarshi24st_cpu.cpp:237), except for the interpolation along the energy
axis.  The profile is a positive weighted sum of interpolated intensities, so
in each energy column its relative difference from the blackbody profile is
bounded by the largest relative interpolation error of the Planck function
over the x = E z / kT the column samples (z the redshift factor); that bound
is computed here independently in numpy for the cubic stencil (one node below
the bracket, start clamped) and the linear bracket.  The (logT, log g) and mu
stencils reproduce constants exactly."""
import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from oracle_binding import Oracle, rel_err

LOGE_MIN, LOGE_MAX, N_E = -1.30, 2.00, 166  # the reduced grid keeps nsx_H_v200804's energy axis
# bounds of the redshift factor z = 1/(delta sqrt(1-u)) on the test star (M 1.4,
# R_eq 12 km, 600 Hz, spot within 0.25 rad of colatitude 20 deg):
# 1/sqrt(1-u) ~ 1.24 and the Doppler factor delta within 1 +- 0.12
Z_RANGE = (1.0, 1.5)


def planck(x):
    return x ** 3 / np.expm1(x)


def interp_error(x, cubic):
    """Relative error of the reference's energy interpolant of planck() at x
    on the uniform logEkT axis (cpu/nsx.hpp:159-167, 252-256; linear:
    cpu/nsx.hpp:304-316)."""
    h = (LOGE_MAX - LOGE_MIN) / (N_E - 1)
    u = np.clip(np.log10(x), LOGE_MIN, LOGE_MAX)
    i = np.clip(np.floor((u - LOGE_MIN) / h).astype(int), 0, N_E - 2)
    nodes = LOGE_MIN + h * np.arange(N_E)
    if cubic:
        s = np.clip(i - 1, 0, N_E - 4)
        v = np.zeros_like(x)
        for k in range(4):
            w = np.ones_like(x)
            for m in range(4):
                if m != k:
                    w *= (u - nodes[s + m]) / (nodes[s + k] - nodes[s + m])
            v += w * planck(10.0 ** nodes[s + k])
    else:
        a = (u - nodes[i]) / h
        v = (1 - a) * planck(10.0 ** nodes[i]) + a * planck(10.0 ** nodes[i + 1])
    return np.abs(v / planck(10.0 ** u) - 1.0)


def column_bounds(energies, kT, cubic):
    out = []
    for E in energies:
        x = np.geomspace(E * Z_RANGE[0] / kT, E * Z_RANGE[1] / kT, 4001)
        out.append(interp_error(x, cubic).max())
    return np.array(out)


def star_pair(mode):
    bb = S.cfg0_star()
    nsx = S.cfg0_star()
    nsx.atmosphere = abi.ATM_NSX
    nsx.nsx_interp = mode
    return bb, nsx


def grid():
    return abi.make_grid(48, 16, sampling=abi.SAMPLE_GL, n_gl=2, energies=list(S.geomspace(0.2, 12.0, 64)))


def check_columns(f, f_bb, bound, slack):
    m = np.abs(f_bb) > 1e-6 * np.abs(f_bb).max()
    r = np.where(m, np.abs(f - f_bb) / np.where(m, np.abs(f_bb), 1.0), 0.0).max(axis=0)
    assert (r <= bound + slack).all(), (r - bound - slack).max()
    return r


@pytest.fixture(scope="module")
def orc(data_root, small_nsx_cfg):
    o = Oracle(data_root, nsx_path=fixtures.ensure_nsx_bb(data_root), responses=("std",), nsx_cfg=small_nsx_cfg)
    yield o
    o.close()


def test_interpolation_bound_is_small_below_a_few_kev():
    # the bound itself: below 1e-5 up to 1 keV at kT 0.35 keV, growing in the
    # Wien tail (x = E z / kT reaches 50 at 12 keV)
    b = column_bounds([0.2, 1.0, 12.0], 0.35, cubic=True)
    assert b[0] < 1e-5 and b[1] < 1e-5 and b[2] > 1e-3


@pytest.mark.parametrize("mode", [abi.NSX_2L2C_LE, abi.NSX_4C, abi.NSX_4C_LE, abi.NSX_4L])
def test_oracle_nsx_path_reproduces_blackbody(orc, mode):
    bb, nsx = star_pair(mode)
    g = grid()
    f_bb, i_bb = orc.profile(bb, g, threads=8)
    f, i = orc.profile(nsx, g, threads=8)
    assert i.n_active == i_bb.n_active
    en = S.geomspace(0.2, 12.0, 64)
    bound = column_bounds(en, bb.spot[0].kT_keV, cubic=mode != abi.NSX_4L)
    r = check_columns(f, f_bb, bound, 1e-7)
    if mode != abi.NSX_4L:
        assert r[en <= 1.0].max() < 1e-5  # the NSX path equals Ibb there
    assert rel_err(f, f_bb) > 0  # the two paths are separate computations
