"""The full-size oracle fixtures (tests/golden/oracle_fullsize) are what the
oracle computes today on the inputs regenerated here: the input hashes, the
stars, and columns of each profile re-run live (the GPU tests compare the
device paths with these fixtures)."""
import os
import sys

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from oracle_binding import Oracle, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_oracle_fullsize as M  # noqa: E402

FIX = os.path.join(HERE, "golden", "oracle_fullsize")


@pytest.fixture(scope="module")
def root():
    return M.ensure_inputs()


def test_inputs_and_stars_are_the_recorded_ones(root):
    want = open(os.path.join(FIX, "INPUTS.sha256")).read().split("\n")
    assert [w for w in want if w] == M.input_hashes(root)
    for name, (_, stars, grid, cols) in M.configs().items():
        fx = np.load(os.path.join(FIX, name + ".npz"))
        assert fx["stars"].tobytes() == M.star_bytes(stars).tobytes(), name
        assert (int(fx["n_side"]), int(fx["n_phase"]), int(fx["response"])) == (grid.n_side, grid.n_phase,
                                                                                  grid.response)
        if cols is not None:
            assert list(fx["cols"]) == cols


@pytest.mark.parametrize("name,star_idx,n_cols", [("cfg0", 0, None), ("cfg1", 0, 2), ("cfg3", 5, 2),
                                                   ("cfg3", 7, 1)])
def test_fixture_columns_rerun_live(root, name, star_idx, n_cols):
    lensing, stars, grid, _ = M.configs()[name]
    fx = np.load(os.path.join(FIX, name + ".npz"))
    orc = Oracle(root, lensing=lensing, nsx_path=fixtures.nsx_path(root, small=False), responses=("std", "std128"))
    e_all, _ = orc.response(grid.response)
    cols = list(fx["cols"])
    pick = list(range(len(cols))) if n_cols is None else [0, len(cols) - 1][:n_cols]
    g = abi.make_grid(grid.n_side, grid.n_phase, sampling=abi.SAMPLE_GL, n_gl=2,
                      energies=[float(e_all[cols[k]]) for k in pick])
    f, info = orc.profile(stars[star_idx], g, threads=8)
    orc.close()
    assert info.n_active == fx["n_active"][star_idx]
    assert rel_err(f, fx["flux"][star_idx][:, pick], floor=0.0) <= 1e-12
