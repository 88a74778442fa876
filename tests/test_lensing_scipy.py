"""The light-bending table generator (gpu_ppm_amd/synth/ppm_synth.c) against
an independent evaluation of the same published steps with the library
routines the reference's c_lensing.py calls: scipy.integrate.quad (QUADPACK
QAGS) for the bending and travel-time integrals (c_lensing.py:47-77,
115-145), numpy.gradient on the non-uniform cos psi nodes (c_lensing.py:100)
and scipy.interpolate.PchipInterpolator (Fritsch-Carlson, c_lensing.py:102-106).
The generator uses adaptive Gauss-Kronrod 7/15, its own non-uniform gradient
and its own PCHIP; agreement to 1e-10 (1e-8 for the lensing factor) pins all
three.  No GPU needed."""
import os
import subprocess

import numpy as np
import pytest

integrate = pytest.importorskip("scipy.integrate")
interpolate = pytest.importorskip("scipy.interpolate")

from gpu_ppm_amd import fixtures

U = (0.3, 0.4, 0.5)
N_COS = 128


def f(x, u, alpha):  # bending integrand, substitution of c_lensing.py:47-54
    s2 = np.sin(alpha) ** 2
    x2 = x * x
    q = (2 - x2 - u * (1 - x2) ** 2 / (1 - u)) * s2
    return x / np.sqrt(1 - s2 + x2 * q)


def g(x, u, alpha):  # travel-time integrand, c_lensing.py:115-122
    s2 = np.sin(alpha) ** 2
    x2 = x * x
    q = (2 - x2 - u * (1 - x2) ** 2 / (1 - u)) * s2
    t = np.sqrt(1 - s2 + x2 * q)
    return x / (t * (1 + t))


def psi_out(u, a):
    return 2 * np.sin(a) / np.sqrt(1 - u) * integrate.quad(f, 0, 1, args=(u, a))[0]


def cdt_out(u, a):
    return 2 * np.sin(a) ** 2 / (1 - u) * integrate.quad(g, 0, 1, args=(u, a))[0]


def p_over_r(u, a):  # periastron of an inward ray, c_lensing.py:65-67
    t = 2 * np.sin(a) / np.sqrt(3 * (1 - u))
    return -t * np.cos((np.arccos(3 * u / t) + 2 * np.pi) / 3)


def psi(u, a):
    if a <= np.pi / 2:
        return psi_out(u, a)
    return 2 * psi_out(u / p_over_r(u, a), np.pi / 2) - psi_out(u, np.pi - a)


def cdt(u, a):
    if a <= np.pi / 2:
        return cdt_out(u, a)
    return 2 * cdt_out(u / p_over_r(u, a), np.pi / 2) - cdt_out(u, np.pi - a)


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("lens"))
    subprocess.run([fixtures.SYNTH, "lensing", d, str(U[0]), str(U[-1]), str(len(U)), str(N_COS)], check=True)
    return {k: np.loadtxt(os.path.join(d, k + ".txt"))
            for k in ("cos_alpha_of_u_cos_psi", "lf_of_u_cos_psi", "cdt_over_R_of_u_cos_alpha")}


def scipy_rows(u):
    cp_out = np.linspace(-0.9, 1.0, N_COS)
    psi_stop = np.arccos(cp_out[0])
    d_alpha = np.pi / (N_COS * 8)
    alphas, psis = [], []
    a = 0.0
    while True:  # alpha march of c_lensing.py:86-94
        alphas.append(a)
        psis.append(psi(u, a))
        a += d_alpha
        if psis[-1] > psi_stop:
            break
    ca, cp = np.cos(np.array(alphas)), np.cos(np.array(psis))
    lf = np.gradient(ca, cp)
    ca_out = interpolate.PchipInterpolator(cp[::-1], ca[::-1])(cp_out)
    lf_out = interpolate.PchipInterpolator(cp[::-1], lf[::-1])(cp_out)
    cdt_row = np.array([cdt(u, np.arccos(c)) for c in np.linspace(-0.2, 1.0, N_COS)])
    return ca_out, lf_out, cdt_row


@pytest.mark.parametrize("k", range(len(U)))
def test_generator_matches_quadpack_pchip_gradient(tables, k):
    ca, lf, cdt_row = scipy_rows(U[k])
    # measured agreement: 1e-13 (cos alpha, c dt / R), 3e-11 (lensing factor)
    np.testing.assert_allclose(tables["cos_alpha_of_u_cos_psi"][k], ca, rtol=0, atol=1e-10)
    np.testing.assert_allclose(tables["lf_of_u_cos_psi"][k], lf, rtol=1e-8, atol=0)
    np.testing.assert_allclose(tables["cdt_over_R_of_u_cos_alpha"][k], cdt_row, rtol=1e-10, atol=1e-12)
