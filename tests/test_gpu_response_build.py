"""The instrument response built on the device (ppm_response_build) equals
the host builder's tables (gpu_ppm_amd/apps/ppm_response.cpp, pinned against
scipy's CubicSpline in test_response_builder.py) on the same synthetic RMF /
ARF / ISM files: E_obs and rsp to 1e-12 of their largest entry."""
import numpy as np
import pytest

from gpu_ppm_amd import lib
from test_response_builder import CH_RANGE, E_RANGE, N_EOBS, instrument, scipy_response  # noqa: F401

pytestmark = pytest.mark.gpu


def test_device_response_equals_host_builder(instrument):  # noqa: F811
    t = instrument
    arf = np.column_stack([t["e_lo"], t["e_hi"], t["area"]])
    ism = np.column_stack([t["ism_e"], t["ism_f"]])
    print("case: response on the device, %d x %d -> %d channels x %d E_obs"
          % (t["rmf"].shape[0], t["rmf"].shape[1], CH_RANGE[1] - CH_RANGE[0], N_EOBS), flush=True)
    e_obs, rsp = lib.response_build(t["rmf"], arf, ism, E_RANGE, CH_RANGE, N_EOBS)
    e_host = np.loadtxt(t["out"] / "E_obs.txt")
    rsp_host = np.loadtxt(t["out"] / "rsp.txt")
    assert rsp.shape == rsp_host.shape
    assert np.abs(e_obs - e_host).max() <= 1e-12 * e_host.max()
    err = np.abs(rsp - rsp_host).max() / np.abs(rsp_host).max()
    print("  device vs host builder: E_obs %.2e, rsp %.2e" % (np.abs(e_obs / e_host - 1).max(), err), flush=True)
    assert err <= 1e-12
    e_ref, rsp_ref = scipy_response(t)
    assert np.abs(rsp - rsp_ref).max() <= 1e-10 * np.abs(rsp_ref).max()
