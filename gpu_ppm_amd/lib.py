"""Python host API over libppm's C ABI.

Mirrors the reference's host interface for this path: the table objects
(Lens, NSX, IsmInst: gpu/lensing.cuh, gpu/nsx.cuh, gpu/ism_inst.cuh) become
one ``Context`` that loads them into HBM once, and ``calc_one_star``
(gpu/devarshi24st_gpu.cu:431-511, gpu/os1_gpu.cu:303-349) becomes
``Context.eval`` (flux and counts for one or many stars).  Errors raise
``PpmError`` carrying the library's status code instead of exiting.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int64, c_uint64, c_void_p

import numpy as np

from . import abi

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libppm.so")
# PPM_LIBRARY=phase selects the build with phase timers in the batched
# integral (ppm_phase_timers); it computes the same results.
PHASE_LIB_PATH = os.path.join(PKG, "libppm_phase.so")


class PpmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("ppm status %d: %s" % (status, message))
        self.status = status


_lib = None


def load():
    """Load libppm.so from the package directory (raises if it is missing:
    there is no fallback implementation)."""
    global _lib
    if _lib is not None:
        return _lib
    path = PHASE_LIB_PATH if os.environ.get("PPM_LIBRARY") == "phase" else LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError("%s is not built; run __graft_entry__.build()" % os.path.basename(path))
    L = ctypes.CDLL(path)
    L.ppm_abi_version.restype = c_int
    L.ppm_strerror.restype = c_char_p
    L.ppm_strerror.argtypes = [c_int]
    L.ppm_last_error.restype = c_char_p
    L.ppm_last_error.argtypes = [c_void_p]
    L.ppm_open.restype = c_int
    L.ppm_open.argtypes = [POINTER(abi.OpenOpts), POINTER(c_void_p)]
    L.ppm_close.restype = c_int
    L.ppm_close.argtypes = [c_void_p]
    L.ppm_response_info.restype = c_int
    L.ppm_response_info.argtypes = [c_void_p, c_int, POINTER(c_int), POINTER(c_int), POINTER(c_double)]
    L.ppm_lensing_domain.restype = c_int
    L.ppm_lensing_domain.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double)]
    L.ppm_eval.restype = c_int
    L.ppm_eval.argtypes = [c_void_p, POINTER(abi.Star), c_int, POINTER(abi.Grid), POINTER(c_double),
                           POINTER(c_double), POINTER(abi.Stats)]
    L.ppm_eval_device.restype = c_int
    L.ppm_eval_device.argtypes = [c_void_p, POINTER(abi.Star), c_int, POINTER(abi.Grid), c_void_p, c_void_p,
                                  POINTER(abi.Stats)]
    L.ppm_eval_device_noalloc.restype = c_int
    L.ppm_eval_device_noalloc.argtypes = [c_void_p, POINTER(abi.Star), POINTER(abi.Grid), c_void_p, c_void_p,
                                          POINTER(abi.Stats)]
    L.ppm_count_active.restype = c_int
    L.ppm_count_active.argtypes = [c_void_p, POINTER(abi.Star), POINTER(abi.Grid), POINTER(c_int64)]
    L.ppm_fold_device.restype = c_int
    L.ppm_fold_device.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p, POINTER(c_double), c_void_p]
    L.ppm_submit.restype = c_int
    L.ppm_submit.argtypes = [c_void_p, POINTER(abi.Star), POINTER(abi.Grid), POINTER(c_uint64)]
    L.ppm_poll.restype = c_int
    L.ppm_poll.argtypes = [c_void_p, c_uint64, POINTER(c_double), POINTER(c_double), POINTER(c_double)]
    L.ppm_cancel.restype = c_int
    L.ppm_cancel.argtypes = [c_void_p, c_uint64]
    L.ppm_serve_stats_get.restype = c_int
    L.ppm_serve_stats_get.argtypes = [c_void_p, POINTER(abi.ServeStats)]
    L.ppm_drain.restype = c_int
    L.ppm_drain.argtypes = [c_void_p]
    L.ppm_eval_fit.restype = c_int
    L.ppm_eval_fit.argtypes = [c_void_p, POINTER(abi.Star), c_int, POINTER(abi.Grid), POINTER(c_double), c_int,
                               POINTER(abi.Fit), POINTER(c_double), POINTER(abi.Stats)]
    L.ppm_tune.restype = c_int
    L.ppm_tune.argtypes = [c_void_p, POINTER(abi.Star), c_int, POINTER(abi.Grid), c_int, POINTER(abi.TuneResult)]
    for f in (L.ppm_set_launch_shape, L.ppm_get_launch_shape):
        f.restype = c_int
        f.argtypes = [c_void_p, POINTER(abi.Grid), c_int, c_int, POINTER(abi.LaunchShape)]
    L.ppm_lensing_build.restype = c_int
    L.ppm_lensing_build.argtypes = [POINTER(abi.LensingSpec), POINTER(c_double), POINTER(c_double),
                                    POINTER(c_double)]
    L.ppm_response_build.restype = c_int
    L.ppm_response_build.argtypes = [POINTER(abi.ResponseSpec), POINTER(c_double), POINTER(c_double)]
    L.ppm_phase_timers.restype = c_int
    L.ppm_phase_timers.argtypes = [c_void_p, POINTER(c_uint64), c_int, c_int]
    _lib = L
    return L


def lensing_build(u_min, u_max, n_u, n_cos, device=0):
    """Lensing tables on the device (ppm_lensing_build): cos(alpha) and the
    lensing factor on the cos(psi) axis, c dt / R on the cos(alpha) axis,
    each [n_u][n_cos]."""
    L = load()
    sp = abi.LensingSpec(u_min, u_max, n_u, n_cos, device, 0)
    ca, lf, dt = (np.zeros((n_u, n_cos)) for _ in range(3))
    st = L.ppm_lensing_build(byref(sp), _dptr(ca), _dptr(lf), _dptr(dt))
    if st != abi.OK:
        raise PpmError(st, L.ppm_strerror(st).decode())
    return ca, lf, dt


def response_build(rmf, arf, ism, e_range, ch_range, n_eobs, device=0):
    """Instrument response on the device (ppm_response_build): rmf
    [n_e_all][n_ch_all], arf [n_e_all][3] (E_lo, E_hi, area), ism [n_ism][2]
    (energy, transmission); returns (E_obs [n_eobs], rsp [n_ch][n_eobs])."""
    L = load()
    rmf = np.ascontiguousarray(rmf, dtype=np.float64)
    arf = np.ascontiguousarray(arf, dtype=np.float64)
    ism = np.ascontiguousarray(ism, dtype=np.float64)
    sp = abi.ResponseSpec(_dptr(rmf), _dptr(arf), _dptr(ism), rmf.shape[0], rmf.shape[1], ism.shape[0],
                          e_range[0], e_range[1], ch_range[0], ch_range[1], n_eobs, device)
    e_obs = np.zeros(n_eobs)
    rsp = np.zeros((ch_range[1] - ch_range[0], n_eobs))
    st = L.ppm_response_build(byref(sp), _dptr(e_obs), _dptr(rsp))
    if st != abi.OK:
        raise PpmError(st, L.ppm_strerror(st).decode())
    return e_obs, rsp


def _dptr(a):
    return a.ctypes.data_as(POINTER(c_double))


def _star_array(stars):
    if isinstance(stars, abi.Star):
        stars = [stars]
    arr = (abi.Star * len(stars))()
    for k, s in enumerate(stars):
        arr[k] = s
    return arr


class Context:
    """Tables resident in HBM plus the evaluation entry points."""

    def __init__(self, root=".", lensing="std", nsx_path=None, responses=(), device=0,
                 nsx_config=None, hbm_budget_bytes=0):
        L = load()
        self._keep = []
        opts = abi.OpenOpts()
        opts.root = os.fsencode(root)
        opts.lensing_setting = lensing.encode()
        # NULL -> default path under root; "" -> no NSX grid
        opts.nsx_path = None if nsx_path is None else os.fsencode(nsx_path)
        opts.response_settings = ",".join(responses).encode()
        if nsx_config is not None:
            self._keep.append(nsx_config)
            opts.nsx_config = ctypes.pointer(nsx_config)
        opts.device = device
        opts.hbm_budget_bytes = hbm_budget_bytes
        h = c_void_p()
        st = L.ppm_open(byref(opts), byref(h))
        if st != abi.OK:
            raise PpmError(st, L.ppm_strerror(st).decode())
        self._h = h
        self.responses = list(responses)

    # -- helpers
    def _check(self, st):
        if st != abi.OK:
            msg = load().ppm_last_error(self._h).decode() or load().ppm_strerror(st).decode()
            raise PpmError(st, msg)

    def close(self):
        if getattr(self, "_h", None):
            load().ppm_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def response_info(self, k):
        n_ch, n_e = c_int(), c_int()
        self._check(load().ppm_response_info(self._h, k, byref(n_ch), byref(n_e), None))
        e = np.zeros(n_e.value)
        self._check(load().ppm_response_info(self._h, k, None, None, _dptr(e)))
        return n_ch.value, e

    def lensing_domain(self):
        a, b, c = c_double(), c_double(), c_double()
        self._check(load().ppm_lensing_domain(self._h, byref(a), byref(b), byref(c)))
        return a.value, b.value, c.value

    def n_energy(self, grid):
        if grid.energies_keV and grid.n_energy > 0:
            return grid.n_energy
        return self.response_info(grid.response)[1].size

    # -- evaluation
    def eval(self, stars, grid, counts=True):
        """Flux [n][n_phase][n_energy] and counts [n][n_ch][n_phase] (None
        without a response) plus a stats struct."""
        arr = _star_array(stars)
        n = len(arr)
        ne = self.n_energy(grid)
        flux = np.zeros((n, grid.n_phase, ne))
        cnt = None
        if counts and grid.response >= 0:
            n_ch = self.response_info(grid.response)[0]
            cnt = np.zeros((n, n_ch, grid.n_phase))
        stats = abi.Stats()
        self._check(load().ppm_eval(self._h, arr, n, byref(grid), _dptr(flux),
                                    _dptr(cnt) if cnt is not None else None, byref(stats)))
        return flux, cnt, stats

    def eval_fit(self, stars, grid, data, counts=False):
        """Fit statistics of the stars' counts against data ([n_ch][n_phase]
        shared, or [n][n_ch][n_phase]); returns (chi2, loglike, model_total)
        arrays of length n, the counts if asked for, and the stats."""
        arr = _star_array(stars)
        n = len(arr)
        d = np.ascontiguousarray(data, dtype=np.float64)
        per_star = 1 if d.ndim == 3 else 0
        fits = (abi.Fit * n)()
        cnt = None
        if counts:
            cnt = np.zeros((n, self.response_info(grid.response)[0], grid.n_phase))
        stats = abi.Stats()
        self._check(load().ppm_eval_fit(self._h, arr, n, byref(grid), _dptr(d), per_star, fits,
                                        _dptr(cnt) if cnt is not None else None, byref(stats)))
        out = np.array([[f.chi2, f.loglike, f.model_total] for f in fits])
        return out[:, 0], out[:, 1], out[:, 2], cnt, stats

    def eval_device(self, stars, grid, d_flux, d_counts=None):
        """Outputs to device pointers (ints, e.g. gpu_ppm_amd.hip.DeviceArray.ptr)."""
        arr = _star_array(stars)
        stats = abi.Stats()
        self._check(load().ppm_eval_device(self._h, arr, len(arr), byref(grid), c_void_p(d_flux),
                                           c_void_p(d_counts) if d_counts else None, byref(stats)))
        return stats

    def eval_device_noalloc(self, star, grid, d_flux, d_counts=None):
        """One star on the plain path with no device allocation (the
        library's fixed scratch); outputs to device pointers."""
        stats = abi.Stats()
        self._check(load().ppm_eval_device_noalloc(self._h, byref(star), byref(grid), c_void_p(d_flux),
                                                   c_void_p(d_counts) if d_counts else None, byref(stats)))
        return stats

    def count_active(self, star, grid):
        n = c_int64()
        self._check(load().ppm_count_active(self._h, byref(star), byref(grid), byref(n)))
        return n.value

    def fold_device(self, response, n_batch, n_phase, d_flux, scale, d_counts):
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        self._check(load().ppm_fold_device(self._h, response, n_batch, n_phase, c_void_p(d_flux),
                                           _dptr(sc), c_void_p(d_counts)))

    # -- launch shape of the batched path
    def tune(self, stars, grid, repeats=0):
        """Autotune the batched launch shape for this grid shape on these
        stars; returns abi.TuneResult (the shape is kept in the context)."""
        arr = _star_array(stars)
        r = abi.TuneResult()
        self._check(load().ppm_tune(self._h, arr, len(arr), byref(grid), repeats, byref(r)))
        return r

    def set_launch_shape(self, grid, n_stars, atmosphere, items_per_slot=0, wg_per_cu=0, energies_per_lane=0):
        sh = abi.LaunchShape(items_per_slot, wg_per_cu, energies_per_lane, 0)
        self._check(load().ppm_set_launch_shape(self._h, byref(grid), n_stars, atmosphere, byref(sh)))

    def launch_shape(self, grid, n_stars, atmosphere):
        """(items_per_slot, wg_per_cu, energies_per_lane) stored for this grid
        shape; (0, 0, 0): none."""
        sh = abi.LaunchShape()
        self._check(load().ppm_get_launch_shape(self._h, byref(grid), n_stars, atmosphere, byref(sh)))
        return sh.items_per_slot, sh.wg_per_cu, sh.energies_per_lane

    def phase_timers(self, reset=False):
        """Wavefront-cycle totals of the batched integral's phases since the
        last reset, as a dict (abi.PHASE_NAMES, plus "wavefronts"); raises
        PpmError (PPM_ERR_UNSUPPORTED) unless the library is the
        phase-timer build (PPM_LIBRARY=phase)."""
        n = len(abi.PHASE_NAMES) + 1
        buf = (c_uint64 * n)()
        self._check(load().ppm_phase_timers(self._h, buf, n, 1 if reset else 0))
        out = dict(zip(abi.PHASE_NAMES, (int(v) for v in buf[:n - 1])))
        out["wavefronts"] = int(buf[n - 1])
        return out

    # -- serving
    def submit(self, star, grid):
        t = c_uint64()
        self._check(load().ppm_submit(self._h, byref(star), byref(grid), byref(t)))
        return t.value

    def poll(self, ticket, grid, out=None):
        """(status, flux, counts, latency_ms); status PENDING until done.
        out: optional (flux, counts) arrays to fill instead of new ones."""
        if out is not None:
            flux, cnt = out
        else:
            ne = self.n_energy(grid)
            flux = np.zeros((grid.n_phase, ne))
            cnt = None
            if grid.response >= 0:
                cnt = np.zeros((self.response_info(grid.response)[0], grid.n_phase))
        lat = c_double()
        st = load().ppm_poll(self._h, ticket, _dptr(flux), _dptr(cnt) if cnt is not None else None,
                             byref(lat))
        if st not in (abi.OK, abi.PENDING, abi.ERR_CANCELLED, abi.ERR_NOT_FOUND):
            self._check(st)
        return st, flux, cnt, lat.value

    def cancel(self, ticket):
        return load().ppm_cancel(self._h, ticket)

    def serve_stats(self):
        out = abi.ServeStats()
        self._check(load().ppm_serve_stats_get(self._h, byref(out)))
        return out

    def drain(self):
        self._check(load().ppm_drain(self._h))
