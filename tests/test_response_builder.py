"""The response builder (gpu_ppm_amd/apps/ppm_response.cpp, a restatement of
c_ism_inst.py:30-101) against an independent evaluation of the same steps
with the library routines the reference calls (numpy.geomspace,
scipy.interpolate.CubicSpline with its not-a-knot default and its exact
piecewise integration), on synthetic RMF / ARF / ISM files in the
reference's text formats.  No GPU needed."""
import os
import subprocess

import numpy as np
import pytest

interpolate = pytest.importorskip("scipy.interpolate")

from gpu_ppm_amd import build

TOOL = os.path.join(build.BIN, "ppm_response")
N_E_ALL, N_CH_ALL = 220, 80
E_RANGE, CH_RANGE, N_EOBS = (10, 200), (5, 70), 64


@pytest.fixture(scope="module")
def instrument(tmp_path_factory):
    d = tmp_path_factory.mktemp("inst")
    rng = np.random.default_rng(3)
    edges = 0.2 + np.cumsum(np.concatenate([[0.0], 0.03 + 0.04 * rng.random(N_E_ALL)]))
    e_lo, e_hi = edges[:-1], edges[1:]
    e_mid = 0.5 * (e_lo + e_hi)
    area = 300 * np.exp(-0.5 * (np.log(e_mid / 1.5) / 0.6) ** 2) * (1 + 0.03 * rng.standard_normal(N_E_ALL))
    ch = np.linspace(0.15, 10.0, N_CH_ALL)
    sig = 0.05 + 0.04 * np.sqrt(e_mid)
    rmf = np.exp(-0.5 * ((ch[None, :] - e_mid[:, None]) / sig[:, None]) ** 2)  # [e][ch]
    rmf /= rmf.sum(axis=1, keepdims=True)
    with open(d / "rmf.txt", "w") as f:
        f.write("# synthetic RMF\n# energy-major\n# e ch value\n")
        for i in range(N_E_ALL):
            for c in range(N_CH_ALL):
                f.write("%d %d %.17g\n" % (i, c, rmf[i, c]))
    with open(d / "arf.txt", "w") as f:
        f.write("# synthetic ARF\n# E_lo E_hi area\n#\n")
        for i in range(N_E_ALL):
            f.write("%.17g %.17g %.17g\n" % (e_lo[i], e_hi[i], area[i]))
    ism_e = np.geomspace(0.1, 15.0, 60)
    ism_f = np.exp(-0.2 * ism_e ** (-8.0 / 3.0))
    np.savetxt(d / "ism.txt", np.column_stack([ism_e, np.zeros_like(ism_e), ism_f]))
    out = d / "out"
    subprocess.run([TOOL, "--rmf", str(d / "rmf.txt"), "--arf", str(d / "arf.txt"), "--ism", str(d / "ism.txt"),
                    "--out", str(out), "--n-eobs", str(N_EOBS), "--e-range", *map(str, E_RANGE),
                    "--ch-range", *map(str, CH_RANGE)], check=True, stdout=subprocess.DEVNULL)
    return dict(out=out, e_lo=e_lo, e_hi=e_hi, area=area, rmf=rmf, ism_e=ism_e, ism_f=ism_f)


def scipy_response(t):
    """The published steps of c_ism_inst.py:43-98 with numpy / scipy."""
    e0, e1 = E_RANGE
    c0, c1 = CH_RANGE
    rmf = t["rmf"].T[c0:c1, e0:e1]                      # [ch][e] after the trim
    lo, hi, area = t["e_lo"][e0:e1], t["e_hi"][e0:e1], t["area"][e0:e1]
    resp = area * rmf
    e_obs = np.geomspace(lo[0], hi[-1], N_EOBS)
    C = np.zeros((len(lo), N_EOBS))
    for i in range(N_EOBS):
        y = np.zeros(N_EOBS)
        y[i] = 1.0
        cs = interpolate.CubicSpline(e_obs, y)
        for j in range(len(lo)):
            C[j, i] = cs.integrate(lo[j], hi[j])
    frac = interpolate.CubicSpline(t["ism_e"], t["ism_f"])(0.5 * (lo + hi)) ** 5
    rsp = (resp * frac) @ C
    rsp[np.abs(rsp) < 1e-300] = 0.0
    return e_obs, rsp


def test_builder_matches_scipy(instrument):
    e_obs, rsp = scipy_response(instrument)
    out = instrument["out"]
    assert open(out / "num.txt").read().split() == [str(CH_RANGE[1] - CH_RANGE[0]), str(N_EOBS)]
    got_e = np.loadtxt(out / "E_obs.txt")
    np.testing.assert_allclose(got_e, e_obs, rtol=1e-14)
    got = np.loadtxt(out / "rsp.txt")
    assert got.shape == rsp.shape
    np.testing.assert_allclose(got, rsp, rtol=0, atol=1e-11 * np.abs(rsp).max())


def test_builder_rejects_inconsistent_files(tmp_path):
    (tmp_path / "arf.txt").write_text("#\n#\n#\n0.1 0.2 10\n0.2 0.3 11\n0.3 0.4 12\n")
    (tmp_path / "rmf.txt").write_text("#\n#\n#\n" + "".join("0 0 %d\n" % k for k in range(7)))
    (tmp_path / "ism.txt").write_text("".join("%g 0 1\n" % e for e in (0.1, 0.2, 0.3, 0.5, 1.0)))
    r = subprocess.run([TOOL, "--rmf", str(tmp_path / "rmf.txt"), "--arf", str(tmp_path / "arf.txt"),
                        "--ism", str(tmp_path / "ism.txt"), "--out", str(tmp_path / "o")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "multiple of the ARF rows" in r.stderr
