"""The count of distinct rotational phases (oracle_binding.distinct_phases),
which pins the number of evaluations of the batched path, checked where it
has a closed form: one phase point (the oracle's active-patch count) and
caps of whole rings (n times the number of distinct offsets frac(p n),
counted by number theory)."""
import math

from gpu_ppm_amd import abi
from oracle_binding import count_active, distinct_phases


def cap_star(theta, phi, rho):
    return abi.make_star([abi.make_spot(theta, rho, 0.3, phi1=phi)], inclination_rad=0.0)


def whole_rings(n_side, rho):
    """Pixel counts of the rings inside a polar cap of radius rho."""
    out = []
    for i in range(1, 2 * n_side):
        z = 1 - i * i / (3.0 * n_side * n_side) if i < n_side else 4.0 / 3 - 2.0 * i / (3 * n_side)
        if math.acos(z) < rho:
            out.append(4 * min(i, n_side))
    return out


def test_one_phase_point_counts_the_active_patches():
    """With one phase point every active patch has its own phase."""
    for n_side, (th, ph, rho) in [(16, (0.6, 0.4, 0.35)), (32, (1.9, 2.0, 0.5)), (64, (0.2, 0.0, 0.15))]:
        st = cap_star(th, ph, rho)
        g = abi.make_grid(n_side, 1, sampling=abi.SAMPLE_POINTS)
        assert distinct_phases(st, g) == count_active(st, g)


def test_uniform_points_on_the_pixel_lattice():
    """Points b / 4 on rings of 4 i pixels: p n = b i is an integer, one
    class, and a whole ring's shifted copies coincide: n phases per ring."""
    n_side, rho = 16, 0.9
    st = cap_star(0.0, 0.0, rho)
    g = abi.make_grid(n_side, 4, sampling=abi.SAMPLE_POINTS)
    rings = whole_rings(n_side, rho)
    assert rings and distinct_phases(st, g) == sum(rings) == count_active(st, g)


def test_gl_points_on_whole_rings():
    """GL points (b + x_g) / 32, x_g = 1/2 -+ 1/(2 sqrt 3): frac(p n) takes
    32 / gcd(n, 32) values per node (b n / 32 mod 1) and the two nodes'
    values differ (x_g n / 32 are irrational apart), so a whole ring of n
    pixels has 2 n 32 / gcd(n, 32) distinct phases."""
    n_side, rho = 32, 1.2
    st = cap_star(0.0, 0.0, rho)
    g = abi.make_grid(n_side, 32, sampling=abi.SAMPLE_GL, n_gl=2)
    rings = whole_rings(n_side, rho)
    expect = sum(n * 2 * 32 // math.gcd(n, 32) for n in rings)
    assert len(rings) > 20 and distinct_phases(st, g) == expect
