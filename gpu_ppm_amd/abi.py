"""ctypes mirror of include/ppm.h (the libppm C ABI).

Only plain data types live here; the library loader is in ``gpu_ppm_amd.lib``.
"""
import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_size_t, c_uint64, c_void_p

PPM_MAX_SPOTS = 4

# ppm_status
OK = 0
ERR_ARG = 1
ERR_IO = 2
ERR_FORMAT = 3
ERR_DEVICE = 4
ERR_NOMEM = 5
ERR_BUSY = 6
ERR_NOT_FOUND = 7
ERR_CANCELLED = 8
PENDING = 9
ERR_UNSUPPORTED = 10

ATM_BLACKBODY = 0
ATM_NSX = 1
NSX_2L2C_LE = 0  # NSX interpolants (ppm_star.nsx_interp)
NSX_4C = 1
NSX_4C_SG = 2
NSX_4C_LE = 3
NSX_4L = 4
BEAM_ISOTROPIC = ord("i")
BEAM_COS2 = ord("c")
BEAM_SIN2 = ord("s")
MEMBER_CAP = 0
MEMBER_DISKDIFF = 1
VARIANT_OS1 = 0
VARIANT_DEVARSHI = 1
SAMPLE_POINTS = 0
SAMPLE_GL = 1
METHOD_AUTO = 0
METHOD_PLAIN = 1
METHOD_BATCHED = 2

# phases of the batched integral reported by ppm_phase_timers, in order (the
# sweep includes the patch set-up and the window ends)
PHASE_NAMES = ("fetch", "ring_setup", "events", "sort", "visible", "split", "sweep", "patch_setup",
               "window_ends", "offsets", "total")
PHASE_COUNT = 12  # PPM_PHASE_COUNT: the phases and the wavefront count

# physical constants shared by the scenario builders (gpu/unit.cuh:3-22)
KB_KEV_PER_K = 8.617333262e-8
KPC_KM = 3.08567758e16


class Spot(ctypes.Structure):
    _fields_ = [
        ("theta1", c_double), ("phi1", c_double), ("angrad1", c_double),
        ("theta2", c_double), ("phi2", c_double), ("angrad2", c_double),
        ("kT_keV", c_double),
        ("membership", c_int), ("beaming", c_int),
    ]


class Star(ctypes.Structure):
    _fields_ = [
        ("mass_msun", c_double), ("radius_eq_km", c_double), ("spin_hz", c_double),
        ("inclination_rad", c_double), ("distance_km", c_double), ("exposure_s", c_double),
        ("logg_override", c_double), ("point_source_below", c_double),
        ("atmosphere", c_int), ("variant", c_int), ("n_spots", c_int), ("nsx_interp", c_int),
        ("spot", Spot * PPM_MAX_SPOTS),
    ]


class Grid(ctypes.Structure):
    _fields_ = [
        ("n_side", c_int), ("n_phase", c_int), ("sampling", c_int), ("n_gl", c_int),
        ("n_fine", c_int), ("n_energy", c_int), ("energies_keV", POINTER(c_double)),
        ("response", c_int), ("method", c_int),
    ]


class NsxConfig(ctypes.Structure):
    _fields_ = [
        ("n_logT", c_int), ("n_logg", c_int), ("n_logEkT", c_int), ("n_mu", c_int),
        ("logT_min", c_double), ("logT_max", c_double),
        ("logg_min", c_double), ("logg_max", c_double),
        ("logEkT_min", c_double), ("logEkT_max", c_double),
    ]


class OpenOpts(ctypes.Structure):
    _fields_ = [
        ("root", c_char_p), ("lensing_setting", c_char_p), ("nsx_path", c_char_p),
        ("response_settings", c_char_p), ("nsx_config", POINTER(NsxConfig)),
        ("device", c_int), ("reserved", c_int), ("hbm_budget_bytes", c_size_t),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", c_double), ("integral_ms", c_double), ("fold_ms", c_double), ("total_ms", c_double),
        ("n_active", c_int64), ("n_eval_ref", c_int64), ("n_eval_done", c_int64),
        ("n_rings", c_int64),
    ]


class Fit(ctypes.Structure):
    _fields_ = [("chi2", c_double), ("loglike", c_double), ("model_total", c_double)]


class LaunchShape(ctypes.Structure):
    _fields_ = [("items_per_slot", c_int), ("wg_per_cu", c_int), ("energies_per_lane", c_int), ("reserved", c_int)]


class TuneResult(ctypes.Structure):
    _fields_ = [("best", LaunchShape), ("best_ms", c_double), ("default_ms", c_double), ("n_tried", c_int),
                ("reserved", c_int)]


class ResponseSpec(ctypes.Structure):
    _fields_ = [("rmf", ctypes.POINTER(ctypes.c_double)), ("arf", ctypes.POINTER(ctypes.c_double)),
                ("ism", ctypes.POINTER(ctypes.c_double)), ("n_e_all", ctypes.c_int), ("n_ch_all", ctypes.c_int),
                ("n_ism", ctypes.c_int), ("e_lo", ctypes.c_int), ("e_hi", ctypes.c_int), ("ch_lo", ctypes.c_int),
                ("ch_hi", ctypes.c_int), ("n_eobs", ctypes.c_int), ("device", ctypes.c_int)]


class LensingSpec(ctypes.Structure):
    _fields_ = [("u_min", ctypes.c_double), ("u_max", ctypes.c_double), ("n_u", ctypes.c_int),
                ("n_cos", ctypes.c_int), ("device", ctypes.c_int), ("reserved", ctypes.c_int)]


class ServeStats(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int64), ("running", ctypes.c_int64), ("launches", ctypes.c_int64 * 2),
                ("overlapped", ctypes.c_int64), ("charged_bytes", ctypes.c_size_t),
                ("budget_bytes", ctypes.c_size_t), ("device_bytes", ctypes.c_size_t),
                ("fixed_bytes", ctypes.c_size_t)]


NSX_H_V200804 = dict(n_logT=35, n_logg=14, n_logEkT=166, n_mu=67, logT_min=5.10, logT_max=6.80,
                     logg_min=13.7, logg_max=15.0, logEkT_min=-1.30, logEkT_max=2.00)
# cpu/nsx.hpp:39-51
NSX_H_V171019 = dict(n_logT=35, n_logg=11, n_logEkT=166, n_mu=67, logT_min=5.10, logT_max=6.80,
                     logg_min=13.7, logg_max=14.7, logEkT_min=-1.30, logEkT_max=2.00)


def nsx_config(**kw):
    d = dict(NSX_H_V200804)
    d.update(kw)
    return NsxConfig(**d)


def make_spot(theta1, angrad1, kT_keV, phi1=0.0, theta2=0.0, phi2=0.0, angrad2=0.0,
              membership=MEMBER_CAP, beaming=BEAM_ISOTROPIC):
    return Spot(theta1=theta1, phi1=phi1, angrad1=angrad1, theta2=theta2, phi2=phi2,
                angrad2=angrad2, kT_keV=kT_keV, membership=membership, beaming=beaming)


def make_star(spots, mass_msun=1.4, radius_eq_km=12.0, spin_hz=600.0, inclination_rad=1.0471975511965976,
              distance_km=KPC_KM, exposure_s=1e6, logg_override=0.0, point_source_below=0.0,
              atmosphere=ATM_BLACKBODY, variant=VARIANT_OS1, nsx_interp=NSX_2L2C_LE):
    st = Star(mass_msun=mass_msun, radius_eq_km=radius_eq_km, spin_hz=spin_hz,
              inclination_rad=inclination_rad, distance_km=distance_km, exposure_s=exposure_s,
              logg_override=logg_override, point_source_below=point_source_below,
              atmosphere=atmosphere, variant=variant, n_spots=len(spots), nsx_interp=nsx_interp)
    if not 1 <= len(spots) <= PPM_MAX_SPOTS:
        raise ValueError("1..%d spots" % PPM_MAX_SPOTS)
    for k, sp in enumerate(spots):
        st.spot[k] = sp
    return st


def make_grid(n_side, n_phase, n_fine=None, sampling=SAMPLE_GL, n_gl=2, energies=None,
              response=-1, method=METHOD_AUTO):
    """energies: sequence of keV (kept alive on the returned object) or None."""
    g = Grid(n_side=n_side, n_phase=n_phase, sampling=sampling, n_gl=n_gl,
             n_fine=n_fine if n_fine else 6 * n_phase, n_energy=0,
             energies_keV=None, response=response, method=method)
    if energies is not None:
        arr = (c_double * len(energies))(*[float(e) for e in energies])
        g.energies_keV = ctypes.cast(arr, POINTER(c_double))
        g.n_energy = len(energies)
        g._keep = arr
    return g


def star_copy(st):
    c = Star()
    ctypes.pointer(c)[0] = st
    return c


__all__ = [n for n in dir() if not n.startswith("_")] + ["c_double", "c_int", "c_int64",
                                                           "c_uint64", "c_void_p", "POINTER"]
