"""Phase timers of the batched integral: the default library has none
(ppm_phase_timers says so), and the phase-timer build (libppm_phase.so,
loaded with PPM_LIBRARY=phase in a child process) computes the same
profiles and reports consistent wavefront-cycle totals."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context, PpmError

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context
root, out = sys.argv[2], sys.argv[3]
c = Context(root, lensing="wide", nsx_path=fixtures.nsx_path(root, small=True), responses=("std", "std128"),
            nsx_config=fixtures.small_nsx_config())
g = abi.make_grid(128, 32, sampling=abi.SAMPLE_GL, n_gl=2, response=1, method=abi.METHOD_BATCHED)
c.phase_timers(reset=True)
f, _, st = c.eval(S.sample_priors(6, seed=11), g)
ph = c.phase_timers()
c.close()
np.save(out + ".npy", f)
json.dump(ph, open(out + ".json", "w"))
print("child: integral %.3f ms, %d wavefronts" % (st.integral_ms, ph["wavefronts"]), flush=True)
"""


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                responses=("std", "std128"), nsx_config=small_nsx_cfg)
    yield c
    c.close()


def test_default_library_reports_no_phase_timers(ctx):
    with pytest.raises(PpmError) as e:
        ctx.phase_timers()
    assert e.value.status == abi.ERR_UNSUPPORTED


def test_phase_timer_build_matches_and_adds_up(ctx, data_root, tmp_path):
    g = abi.make_grid(128, 32, sampling=abi.SAMPLE_GL, n_gl=2, response=1, method=abi.METHOD_BATCHED)
    f_ref, _, _ = ctx.eval(S.sample_priors(6, seed=11), g)
    out = str(tmp_path / "phase")
    print("case: 6 stars N 128 in a child process on libppm_phase.so", flush=True)
    env = dict(os.environ, PPM_LIBRARY="phase")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, data_root, out], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    f = np.load(out + ".npy")
    ph = json.load(open(out + ".json"))
    # the timers only read the clock: the same sums in the same order
    assert np.array_equal(f, f_ref) or np.max(np.abs(f - f_ref)) <= 1e-12 * np.max(np.abs(f_ref))
    assert ph["wavefronts"] > 0 and ph["total"] > 0
    top = ["fetch", "ring_setup", "events", "sort", "visible", "split", "sweep", "offsets"]
    assert sum(ph[k] for k in top) <= ph["total"]
    assert ph["patch_setup"] + ph["window_ends"] <= ph["sweep"]
    assert ph["sweep"] > 0 and ph["patch_setup"] > 0
