"""The FP64 oracle against the reference's own CPU programs.

tests/golden/ref_cpu holds the outputs of cpu/os1_cpu.cpp and
cpu/devarshi24st_cpu.cpp, built from the reference sources by
oracle/Makefile and run on seeded synthetic inputs
(tests/golden/make_ref_goldens.sh).  The reference writes 10 significant
digits, so agreement is limited to ~5e-10 relative.
"""
import os

import numpy as np
import pytest

from gpu_ppm_amd import scenarios as S
from oracle_binding import Oracle, rel_err

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_cpu")
PRINT_TOL = 1e-9  # 10 printed significant digits


@pytest.fixture(scope="module")
def oracle(golden_root):
    o = Oracle(golden_root, nsx_path=os.path.join(golden_root, "ext", "model_data", "nsx_H_v200804.out"),
               responses=("std",))
    yield o
    o.close()


@pytest.mark.parametrize("key", list("abcdefghij"))
def test_os1_light_curves(oracle, key):
    """os1a..j: blackbody spot, cap or point source, 128 phase points at 1 keV
    (cpu/os1_cpu.cpp:17-36, 185-191)."""
    flux, info = oracle.profile(S.os1_star(key), S.os1_grid(), threads=8)
    ref = np.loadtxt(os.path.join(GOLD, "os1%s_cpu.txt" % key))
    assert ref.shape == (128,)
    assert rel_err(flux[:, 0], ref) < PRINT_TOL
    if S.OS1[key][3] < 0.1:
        assert info.n_active == 1  # point-source shortcut (cpu/os1_cpu.cpp:40-42)


@pytest.mark.parametrize("name", ["Ring_Eq", "Crescent_Eq", "Crescent_Polar"])
def test_devarshi_counts_std(oracle, name):
    """Ring / crescent spots, NSX atmosphere, 32 bins x 2 GL, fold to counts
    (cpu/devarshi24st_cpu.cpp:214-258) at the programs' std N_side."""
    n = S.DEVARSHI_N_SIDE[(name, "std")]
    flux, info = oracle.profile(S.devarshi_star(name), S.devarshi_grid(n, response=0), threads=8)
    counts = oracle.fold(flux, 0, S.DEVARSHI_EXPOSURE / S.DEVARSHI_N_PHASE)
    ref = np.loadtxt(os.path.join(GOLD, "%s_std_counts_cpu.txt.gz" % name))
    assert ref.shape == counts.shape == (270, 32)
    assert rel_err(counts, ref) < PRINT_TOL
    assert info.n_clamped > 0  # 12 keV at logT 6 leaves the logEkT axis: clamped like the CPU


@pytest.mark.slow
@pytest.mark.parametrize("name,setting", [("Ring_Polar", "std"), ("Ring_Eq", "high"),
                                          ("Ring_Polar", "high"), ("Crescent_Eq", "high"),
                                          ("Crescent_Polar", "high")])
def test_devarshi_counts_other(oracle, name, setting):
    n = S.DEVARSHI_N_SIDE[(name, setting)]
    flux, _ = oracle.profile(S.devarshi_star(name), S.devarshi_grid(n, response=0), threads=8)
    counts = oracle.fold(flux, 0, S.DEVARSHI_EXPOSURE / S.DEVARSHI_N_PHASE)
    ref = np.loadtxt(os.path.join(GOLD, "%s_%s_counts_cpu.txt.gz" % (name, setting)))
    assert rel_err(counts, ref) < PRINT_TOL


def test_planted_error_is_caught(oracle):
    """A plausible mistake in the oracle's inputs (inclination off by 1e-3
    rad) must fail the golden comparison."""
    st = S.os1_star("f")
    st.inclination_rad += 1e-3
    flux, _ = oracle.profile(st, S.os1_grid(), threads=8)
    ref = np.loadtxt(os.path.join(GOLD, "os1f_cpu.txt"))
    assert rel_err(flux[:, 0], ref) > 1e-5
