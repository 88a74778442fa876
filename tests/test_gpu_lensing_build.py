"""Lensing tables built on the device (ppm_lensing_build) equal the host
generator's (ppm_synth lensing, pinned against scipy's QUADPACK quad,
numpy.gradient and PchipInterpolator in test_lensing_scipy.py) for the std
and wide settings: cos(alpha) and c dt / R to 1e-10, the lensing factor (a
numerical derivative) to 1e-9, relative to each table's largest entry."""
import os
import time

import numpy as np
import pytest

from gpu_ppm_amd import fixtures, lib

pytestmark = pytest.mark.gpu

SETTINGS = {"std": (0.3, 0.5, 256, 512), "wide": (0.15, 0.55, 256, 512)}


def host_tables(root, setting):
    d = fixtures.ensure_lensing(root, setting)
    return [np.loadtxt(os.path.join(d, f)) for f in
            ("cos_alpha_of_u_cos_psi.txt", "lf_of_u_cos_psi.txt", "cdt_over_R_of_u_cos_alpha.txt")]


@pytest.mark.parametrize("setting", ["std", "wide"])
def test_device_tables_equal_host_generator(data_root, setting):
    u0, u1, n_u, n_cos = SETTINGS[setting]
    print("case: lensing tables on the device, %s (u %.2f-%.2f, %d x %d)" % (setting, u0, u1, n_u, n_cos),
          flush=True)
    t0 = time.perf_counter()
    ca, lf, dt = lib.lensing_build(u0, u1, n_u, n_cos)
    t = time.perf_counter() - t0
    h_ca, h_lf, h_dt = host_tables(data_root, setting)
    errs = [np.abs(a - b).max() / np.abs(b).max() for a, b in ((ca, h_ca), (lf, h_lf), (dt, h_dt))]
    print("  %.2f s; cos alpha %.2e, lensing factor %.2e, c dt / R %.2e" % (t, *errs), flush=True)
    assert errs[0] <= 1e-10 and errs[1] <= 1e-9 and errs[2] <= 1e-10
