import hashlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

from gpu_ppm_amd import fixtures  # noqa: E402

DATA = os.path.join(TESTS, "_data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: takes more than ~30 s on the CPU")


@pytest.fixture(scope="session")
def data_root():
    """Synthetic tables (lensing std/wide, responses, reduced NSX grid)."""
    root = os.path.join(DATA, "root")
    fixtures.ensure_all(root, small_nsx=True)
    return root


@pytest.fixture(scope="session")
def small_nsx_cfg():
    return fixtures.small_nsx_config()


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


@pytest.fixture(scope="session")
def golden_root():
    """Inputs the reference CPU programs were run on for tests/golden/ref_cpu,
    regenerated with the same seeds and checked against the recorded hashes."""
    root = os.path.join(DATA, "golden_inputs")
    fixtures.ensure_lensing(root, "std")
    d = os.path.join(root, "tables", "ism_inst", "std")
    fixtures._publish(d, lambda t: fixtures._synth("response", t, 270, 128, 0))
    fixtures.ensure_nsx(root, small=False)
    with open(os.path.join(TESTS, "golden", "ref_cpu", "INPUTS.sha256")) as f:
        for line in f:
            digest, rel = line.split()
            assert sha256(os.path.join(root, rel)) == digest, "generated input differs: " + rel
    return root
