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

from gpu_ppm_amd import abi
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


# ------------------------------------------------------------- t_interp
def _tinterp_ref(name):
    path = os.path.join(GOLD, "t_interp_%s_sub.txt.gz" % name)
    import gzip
    with gzip.open(path, "rt") as f:
        rows = [int(v) for v in f.readline().split()[2:]]
        cols = [int(v) for v in f.readline().split()[2:]]
    return np.loadtxt(path), rows, cols


@pytest.fixture(scope="module")
def tinterp_flux(oracle):
    """Oracle flux of every t_interp star at the sampled energies."""
    _, _, cols = _tinterp_ref("test10")
    en = S.tinterp_energies()[cols]
    out = {}
    for name in S.TINTERP:
        flux, _ = oracle.profile(S.tinterp_star(name), S.tinterp_grid(energies=en), threads=8)
        out[name] = flux
    return out


@pytest.mark.parametrize("name", sorted(S.TINTERP))
def test_tinterp_interpolants(tinterp_flux, name):
    """cpu/t_interp.cpp: one element at logT 6 through Interp_IT3_4c,
    4c with use_sg, 4c_le and 2l2c_le (interp_type 0..3), 256 bins x 2 GL
    nodes, 1024 fine phases; sampled rows and energies of its output."""
    ref, rows, _ = _tinterp_ref(name)
    flux = tinterp_flux[name][rows]
    assert ref.shape == flux.shape == (33, 65)
    assert rel_err(flux, ref) < PRINT_TOL
    # below the comparison floor: the sign of the cubic's Wien-tail
    # undershoot, and use_sg's zeros, agree entry by entry
    mode = S.TINTERP_MODE[S.TINTERP[name][2]]
    if mode == abi.NSX_4C_SG:
        assert (ref >= 0).all() and (flux[ref == 0] == 0).all() and (flux >= 0).all()
    else:
        assert (flux[ref < 0] < 0).all() and (ref < 0).any()


def test_tinterp_wrong_interpolant_is_caught(oracle):
    """Each golden pins its own interpolant: evaluating test10 (4c) with
    2l2c_le or 4l, or test12 (4c_le) with 4l, misses the reference by far
    more than the print precision."""
    ref, rows, cols = _tinterp_ref("test10")
    en = S.tinterp_energies()[cols]
    for mode in (abi.NSX_2L2C_LE, abi.NSX_4L):
        flux, _ = oracle.profile(S.tinterp_star("test10", nsx_interp=mode), S.tinterp_grid(energies=en), threads=8)
        assert rel_err(flux[rows], ref) > 1e-6
