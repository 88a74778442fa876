"""Synthetic inputs are written in the reference's formats and are
reproducible from their seeds."""
import os
import subprocess

import numpy as np

from gpu_ppm_amd import fixtures
from gpu_ppm_amd import scenarios as S


def test_lensing_format(data_root):
    d = os.path.join(data_root, "tables", "lensing", "std")
    assert open(os.path.join(d, "u.txt")).read().split() == ["0.29999999999999999", "0.5", "256"]
    for name, shape in (("cos_alpha_of_u_cos_psi.txt", (256, 512)), ("lf_of_u_cos_psi.txt", (256, 512)),
                        ("cdt_over_R_of_u_cos_alpha.txt", (256, 512))):
        assert np.loadtxt(os.path.join(d, name)).shape == shape


def test_response_format_and_determinism(data_root, tmp_path):
    d = os.path.join(data_root, "tables", "ism_inst", "std")
    assert open(os.path.join(d, "num.txt")).read().split() == ["300", "360"]
    e = np.loadtxt(os.path.join(d, "E_obs.txt"))
    assert e.shape == (360,) and e[0] == 0.2 and e[-1] == 12.0
    assert np.allclose(e, S.geomspace(0.2, 12.0, 360), rtol=1e-14)
    rsp = np.loadtxt(os.path.join(d, "rsp.txt"))
    assert rsp.shape == (300, 360) and np.all(rsp >= 0)
    # the effective area times ISM peaks in the soft band
    assert 0.8 < e[np.argmax(rsp.sum(axis=0))] < 3.0
    subprocess.run([fixtures.SYNTH, "response", str(tmp_path), "300", "360", "0"], check=True)
    assert open(os.path.join(d, "rsp.txt")).read() == open(tmp_path / "rsp.txt").read()


def test_small_nsx_grid(data_root):
    p = fixtures.nsx_path(data_root, small=True)
    a = np.loadtxt(p, max_rows=67 * 3)
    nT, ng, nE, nm = fixtures.SMALL_NSX_DIMS
    assert a.shape[1] == 5
    mu = a[:nm, 1]
    assert mu[0] == 1.0 and np.all(np.diff(mu) < 0) and mu[-1] > 0
    assert np.allclose(a[:nm, 0], -1.3) and np.allclose(a[nm:2 * nm, 0], -1.28)
    # limb-darkened Planck shape: intensity falls toward the limb
    assert np.all(np.diff(a[:nm, 2]) < 0.01)
