"""Scenario builders: the reference programs' hard-coded stars and the
benchmark configurations, expressed as ppm_star / ppm_grid values.

Reference scenarios:
  os1 a..j        cpu/os1_cpu.cpp:17-36 (N_side 256, 128 phase points, 1 keV)
  devarshi stars  cpu/devarshi24st_cpu.cpp:17-22, :48-59, :267-285
  t_interp tests  cpu/t_interp.cpp:19-33, :242-261 (NSX interpolants)
"""
import math

import numpy as np

from . import abi

DEG = math.pi / 180.0


# ------------------------------------------------------------------ os1
OS1 = {  # name: (nu, spot colatitude deg, inclination deg, angular radius, beaming)
    "a": (600, 90, 90, 0.01, "i"),
    "b": (600, 90, 90, 1.0, "i"),
    "c": (200, 90, 90, 0.01, "i"),
    "d": (1, 90, 90, 1.0, "i"),
    "e": (600, 60, 30, 1.0, "i"),
    "f": (600, 20, 80, 1.0, "i"),
    "g": (600, 60, 30, 1.0, "c"),
    "h": (600, 60, 30, 1.0, "s"),
    "i": (600, 20, 80, 1.0, "c"),
    "j": (600, 20, 80, 1.0, "s"),
}
OS1_N_SIDE = 256
OS1_N_PHASE = 128
OS1_E_KEV = 1.0


def os1_star(key):
    nu, th, inc, rho, beam = OS1[key]
    spot = abi.make_spot(th * DEG, rho, 0.35, membership=abi.MEMBER_CAP, beaming=ord(beam))
    return abi.make_star([spot], mass_msun=1.4, radius_eq_km=12.0, spin_hz=nu,
                         inclination_rad=inc * DEG, distance_km=0.2 * abi.KPC_KM,
                         point_source_below=0.1, atmosphere=abi.ATM_BLACKBODY,
                         variant=abi.VARIANT_OS1)


def os1_grid(n_side=OS1_N_SIDE, method=abi.METHOD_AUTO):
    return abi.make_grid(n_side, OS1_N_PHASE, n_fine=3 * OS1_N_PHASE, sampling=abi.SAMPLE_POINTS,
                         n_gl=1, energies=[OS1_E_KEV], response=-1, method=method)


# ------------------------------------------------------------- devarshi
DEVARSHI = {  # name: (theta1, phi1, angrad1, theta2, phi2, angrad2)
    "Ring_Eq": (0.5 * math.pi, -0.2, 0.5, 0.5 * math.pi, 0.0, 0.25),
    "Ring_Polar": (math.pi - 0.001, -0.02, 0.05, math.pi - 0.001, 0.0, 0.025),
    "Crescent_Eq": (1.0, -1.0, 1.4, 0.5, 0.0, 0.5 * math.pi - 0.001),
    "Crescent_Polar": (1.15, -0.1, 1.3, 1.75, 0.0, 0.5 * math.pi - 0.001),
}
DEVARSHI_N_SIDE = {  # (name, setting) -> N_side, cpu/devarshi24st_cpu.cpp:276-285
    ("Ring_Eq", "std"): 173, ("Ring_Polar", "std"): 1708, ("Crescent_Eq", "std"): 128,
    ("Crescent_Polar", "std"): 189, ("Ring_Eq", "high"): 345, ("Ring_Polar", "high"): 3413,
    ("Crescent_Eq", "high"): 247, ("Crescent_Polar", "high"): 382,
}
DEVARSHI_KT_KEV = 1e6 * abi.KB_KEV_PER_K  # logT = 6.0
DEVARSHI_EXPOSURE = 1e6
DEVARSHI_N_PHASE = 32


def devarshi_star(name, kT_keV=DEVARSHI_KT_KEV):
    t1, p1, r1, t2, p2, r2 = DEVARSHI[name]
    spot = abi.make_spot(t1, r1, kT_keV, phi1=p1, theta2=t2, phi2=p2, angrad2=r2,
                         membership=abi.MEMBER_DISKDIFF)
    return abi.make_star([spot], mass_msun=1.4, radius_eq_km=12.0, spin_hz=300.0,
                         inclination_rad=2.391, distance_km=0.15 * abi.KPC_KM,
                         exposure_s=DEVARSHI_EXPOSURE, atmosphere=abi.ATM_NSX,
                         variant=abi.VARIANT_DEVARSHI)


def devarshi_grid(n_side, response=0, n_phase=DEVARSHI_N_PHASE, method=abi.METHOD_AUTO):
    return abi.make_grid(n_side, n_phase, n_fine=6 * n_phase, sampling=abi.SAMPLE_GL, n_gl=2,
                         energies=None, response=response, method=method)


# ------------------------------------------------------------- t_interp
TINTERP = {  # name: (inclination, spot colatitude, interp_type), cpu/t_interp.cpp:248-261
    "test10": (0.01, 1.983, 0), "test11": (0.01, 1.983, 1),
    "test12": (0.01, 1.983, 2), "test13": (0.01, 1.983, 3),
    "test20": (0.964, math.pi - 0.01, 0), "test21": (0.964, math.pi - 0.01, 1),
    "test22": (0.964, math.pi - 0.01, 2), "test23": (0.964, math.pi - 0.01, 3),
}
# interp_type -> interpolant (cpu/t_interp.cpp:210-219)
TINTERP_MODE = {0: abi.NSX_4C, 1: abi.NSX_4C_SG, 2: abi.NSX_4C_LE, 3: abi.NSX_2L2C_LE}
TINTERP_N_PHASE = 256
TINTERP_N_FINE = 1024


def ref_linspace(beg, end, n):
    """linspace of cpu/ut.hpp:7-14 (beg + i * step)."""
    step = (end - beg) / (n - 1)
    return np.array([beg + i * step for i in range(n)])


def tinterp_energies():
    return ref_linspace(0.1, 5.0, 1024)  # cpu/t_interp.cpp:33


def tinterp_star(name, nsx_interp=None):
    """One 0.01 rad spot (a point-source element, cpu/grid.hpp:17-27) at
    logT 6, M 1.4, R 12 km, 600 Hz, D 0.15 kpc; log g from the spot's
    colatitude (cpu/t_interp.cpp:20-58, 76-97)."""
    inc, th, it = TINTERP[name]
    spot = abi.make_spot(th, 0.01, DEVARSHI_KT_KEV, membership=abi.MEMBER_CAP)
    return abi.make_star([spot], mass_msun=1.4, radius_eq_km=12.0, spin_hz=600.0, inclination_rad=inc,
                         distance_km=0.15 * abi.KPC_KM, point_source_below=0.1, atmosphere=abi.ATM_NSX,
                         variant=abi.VARIANT_DEVARSHI,
                         nsx_interp=TINTERP_MODE[it] if nsx_interp is None else nsx_interp)


def tinterp_grid(energies=None, method=abi.METHOD_AUTO):
    """256 bins x 2 GL nodes, 1024 fine phases (cpu/t_interp.cpp:26, 245-246, 188-196)."""
    return abi.make_grid(16, TINTERP_N_PHASE, n_fine=TINTERP_N_FINE, sampling=abi.SAMPLE_GL, n_gl=2,
                         energies=tinterp_energies() if energies is None else energies, response=-1,
                         method=method)


# ------------------------------------------------- benchmark configurations
def geomspace(lo, hi, n):
    """numpy.geomspace with exact end points (c_ism_inst.py:61)."""
    e = np.exp(np.log(lo) + np.arange(n) * (np.log(hi) - np.log(lo)) / (n - 1))
    e[0], e[-1] = lo, hi
    return e


def cfg0_star():
    """Single circular spot, blackbody kT 0.35 keV, os1 code path."""
    spot = abi.make_spot(20 * DEG, 0.25, 0.35, membership=abi.MEMBER_CAP)
    return abi.make_star([spot], mass_msun=1.4, radius_eq_km=12.0, spin_hz=600.0,
                         inclination_rad=60 * DEG, distance_km=abi.KPC_KM, exposure_s=1e6,
                         atmosphere=abi.ATM_BLACKBODY, variant=abi.VARIANT_OS1)


def cfg1_star():
    """The same oblate star with the NSX-format atmosphere (logT from kT)."""
    st = cfg0_star()
    st.atmosphere = abi.ATM_NSX
    return st


def cfg2_star():
    """Ring 0.15-0.4 rad at 35 deg (kT 0.3) + crescent 0.5 minus an offset 0.35
    rad disk at 140 deg (kT 0.25), log g 14.2."""
    ring = abi.make_spot(35 * DEG, 0.4, 0.3, theta2=35 * DEG, angrad2=0.15,
                         membership=abi.MEMBER_DISKDIFF)
    cres = abi.make_spot(140 * DEG, 0.5, 0.25, theta2=140 * DEG + 0.1, angrad2=0.35,
                         membership=abi.MEMBER_DISKDIFF)
    return abi.make_star([ring, cres], mass_msun=1.4, radius_eq_km=12.0, spin_hz=300.0,
                         inclination_rad=2.391, distance_km=0.15 * abi.KPC_KM, exposure_s=1e6,
                         logg_override=14.2, atmosphere=abi.ATM_NSX, variant=abi.VARIANT_DEVARSHI)


PRIORS = dict(mass=(1.0, 2.2), radius=(9.0, 15.0), spin=(200.0, 700.0), inc=(5 * DEG, 90 * DEG),
              colat=(0.0, math.pi), rho=(0.02, 1.2), kT=(0.1, 0.5))
RS_SUN_KM = 2.9532501


def polar_u(mass, radius, spin):
    """u = R_s / R at the pole of the oblate star (the largest u on the star)."""
    rs = mass * RS_SUN_KM
    ob = 2 * math.pi * spin * math.sqrt(radius ** 3 / mass) * 2.745011592867327e-6
    o2 = ob * ob * (-0.788 + 1.030 * 0.5 * rs / radius)
    return rs / (radius * (1.0 + o2))


def sample_priors(n, seed=0, u_max=0.55, atmosphere=abi.ATM_NSX):
    """n stars from the batched-evaluation priors, rejection-sampled so the
    whole surface stays inside the lensing table's u axis (u <= u_max)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        m = rng.uniform(*PRIORS["mass"])
        r = rng.uniform(*PRIORS["radius"])
        nu = rng.uniform(*PRIORS["spin"])
        inc = rng.uniform(*PRIORS["inc"])
        th = rng.uniform(*PRIORS["colat"])
        rho = rng.uniform(*PRIORS["rho"])
        kT = rng.uniform(*PRIORS["kT"])
        if polar_u(m, r, nu) > u_max:
            continue
        spot = abi.make_spot(th, rho, kT, membership=abi.MEMBER_CAP)
        out.append(abi.make_star([spot], mass_msun=m, radius_eq_km=r, spin_hz=nu,
                                 inclination_rad=inc, distance_km=abi.KPC_KM, exposure_s=1e6,
                                 atmosphere=atmosphere, variant=abi.VARIANT_OS1))
    return out


class PriorWalk:
    """Seeded reflecting random walk over the batched-evaluation priors (the
    serving workload): each step moves every parameter by a uniform amount
    within +-`step` of its range, reflecting at the bounds; steps that would
    leave the lensing table's u axis are redrawn."""

    KEYS = ("mass", "radius", "spin", "inc", "colat", "rho", "kT")

    def __init__(self, seed=0, step=0.02, u_max=0.55, atmosphere=abi.ATM_NSX):
        self.rng = np.random.default_rng(seed)
        self.step, self.u_max, self.atmosphere = step, u_max, atmosphere
        while True:
            x = {k: self.rng.uniform(*PRIORS[k]) for k in self.KEYS}
            if polar_u(x["mass"], x["radius"], x["spin"]) <= u_max:
                break
        self.x = x

    @staticmethod
    def _reflect(v, lo, hi):
        w = hi - lo
        v = (v - lo) % (2 * w)
        return lo + (v if v <= w else 2 * w - v)

    def next(self):
        while True:
            y = {}
            for k in self.KEYS:
                lo, hi = PRIORS[k]
                y[k] = self._reflect(self.x[k] + self.rng.uniform(-self.step, self.step) * (hi - lo), lo, hi)
            if polar_u(y["mass"], y["radius"], y["spin"]) <= self.u_max:
                break
        self.x = y
        spot = abi.make_spot(y["colat"], y["rho"], y["kT"], membership=abi.MEMBER_CAP)
        return abi.make_star([spot], mass_msun=y["mass"], radius_eq_km=y["radius"], spin_hz=y["spin"],
                             inclination_rad=y["inc"], distance_km=abi.KPC_KM, exposure_s=1e6,
                             atmosphere=self.atmosphere, variant=abi.VARIANT_OS1)
