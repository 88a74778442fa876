"""ctypes binding of the FP64 oracle (oracle/liboracle.so) for the tests.

Test infrastructure: the oracle is the checker the GPU library is compared
against, never part of the product path.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int64, c_void_p

import numpy as np

from gpu_ppm_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


class OrInfo(ctypes.Structure):
    _fields_ = [("n_active", c_int64), ("n_eval_ref", c_int64), ("n_rings", c_int64),
                ("n_rings_kept", c_int64), ("n_clamped", c_int64)]


class OrResponse(ctypes.Structure):
    _fields_ = [("n_ch", c_int), ("n_e", c_int), ("e_obs", POINTER(c_double)),
                ("rsp", POINTER(c_double))]


class OrNsx(ctypes.Structure):
    _fields_ = [("cfg", abi.NsxConfig), ("mu", POINTER(c_double)), ("data", POINTER(c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.or_lensing_load.restype = c_void_p
        L.or_lensing_load.argtypes = [c_char_p, c_char_p, c_int]
        L.or_nsx_load.restype = POINTER(OrNsx)
        L.or_nsx_load.argtypes = [c_char_p, POINTER(abi.NsxConfig), c_char_p, c_int]
        L.or_response_load.restype = POINTER(OrResponse)
        L.or_response_load.argtypes = [c_char_p, c_char_p, c_int]
        L.or_lensing_free.argtypes = [c_void_p]
        L.or_nsx_free.argtypes = [c_void_p]
        L.or_response_free.argtypes = [c_void_p]
        L.or_profile.restype = c_int
        L.or_profile.argtypes = [c_void_p, c_void_p, c_void_p, POINTER(abi.Star), POINTER(abi.Grid),
                                 POINTER(c_double), POINTER(c_double), POINTER(OrInfo), c_int,
                                 c_char_p, c_int]
        L.or_count_active.restype = c_int64
        L.or_count_active.argtypes = [POINTER(abi.Star), POINTER(abi.Grid)]
        L.or_fold.argtypes = [c_void_p, POINTER(c_double), c_int, c_double, POINTER(c_double)]
        L.or_healpix_ring.argtypes = [c_int, c_int, POINTER(c_double), POINTER(c_int),
                                      POINTER(c_double)]
        L.or_healpix_phi.restype = c_double
        L.or_healpix_phi.argtypes = [c_int, c_int, c_int]
        L.or_nsx_interp.restype = c_double
        L.or_nsx_interp.argtypes = [c_void_p, c_int, c_double, c_double, c_double, c_double]
        L.or_lensing_lookup.argtypes = [c_void_p, c_double, c_double, POINTER(c_double),
                                        POINTER(c_double), POINTER(c_double)]
        L.or_gl_node.restype = c_double
        L.or_gl_node.argtypes = [c_int, c_int, c_double, c_double]
        L.or_gl_weight.restype = c_double
        L.or_gl_weight.argtypes = [c_int, c_int, c_double]
        L.or_vec_hunt.restype = c_int
        L.or_vec_hunt.argtypes = [POINTER(c_double), c_int, c_double, POINTER(c_double)]
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(POINTER(c_double))


class Oracle:
    """Loaded tables + profile evaluation in FP64."""

    def __init__(self, root, lensing="std", nsx_path=None, responses=(), nsx_cfg=None):
        L = lib()
        err = ctypes.create_string_buffer(512)
        self.lt = L.or_lensing_load(os.path.join(root, "tables", "lensing", lensing).encode(), err, 512)
        if not self.lt:
            raise IOError(err.value.decode())
        self.nsx = None
        if nsx_path:
            cfg = nsx_cfg if nsx_cfg is not None else abi.nsx_config()
            self.nsx = L.or_nsx_load(nsx_path.encode(), byref(cfg), err, 512)
            if not self.nsx:
                raise IOError(err.value.decode())
        self.responses = []
        for s in responses:
            r = L.or_response_load(os.path.join(root, "tables", "ism_inst", s).encode(), err, 512)
            if not r:
                raise IOError(err.value.decode())
            self.responses.append(r)

    def response(self, k):
        r = self.responses[k].contents
        e = np.ctypeslib.as_array(r.e_obs, shape=(r.n_e,)).copy()
        m = np.ctypeslib.as_array(r.rsp, shape=(r.n_ch, r.n_e)).copy()
        return e, m

    def profile(self, star, grid, threads=1):
        L = lib()
        rsp = self.responses[grid.response] if grid.response >= 0 else None
        if grid.energies_keV:
            ne = grid.n_energy
            en = None
        else:
            ne = rsp.contents.n_e
            en = None
        flux = np.zeros((grid.n_phase, ne))
        info = OrInfo()
        err = ctypes.create_string_buffer(512)
        rc = L.or_profile(self.lt, self.nsx, rsp, byref(star), byref(grid), en if en else grid.energies_keV,
                          _dptr(flux), byref(info), threads, err, 512)
        if rc:
            raise RuntimeError(err.value.decode())
        return flux, info

    def fold(self, flux, response, scale):
        L = lib()
        r = self.responses[response].contents
        counts = np.zeros((r.n_ch, flux.shape[0]))
        f = np.ascontiguousarray(flux, dtype=np.float64)
        L.or_fold(self.responses[response], _dptr(f), flux.shape[0], scale, _dptr(counts))
        return counts

    def close(self):
        L = lib()
        if self.lt:
            L.or_lensing_free(self.lt)
            self.lt = None
        if self.nsx:
            L.or_nsx_free(self.nsx)
            self.nsx = None
        for r in self.responses:
            L.or_response_free(r)
        self.responses = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def count_active(star, grid):
    return lib().or_count_active(byref(star), byref(grid))


def rel_err(x, ref, floor=1e-6):
    """max |x - ref| / |ref| over entries with |ref| > floor * max|ref|."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    m = np.abs(ref) > floor * np.abs(ref).max()
    if not m.any():
        return 0.0
    return float(np.max(np.abs(x[m] - ref[m]) / np.abs(ref[m])))


def distinct_phases(star, grid):
    """Distinct rotational phases among the reference's (phase point, active
    patch) pairs, summed over rings and spots.

    The reference evaluates every phase point p at every active patch j of a
    ring at the rotational phase frac(p + (j - s/2) / n)
    (cpu/devarshi24st_cpu.cpp:214-239); pairs of one ring with equal phases
    need one evaluation.  Two pairs share a phase exactly when their
    fractional offsets frac(p n) agree and (floor(p n) + j) mod n agree; the
    offsets are compared within 1e-9 and floor(p n) is taken up when the
    fraction exceeds 1 - 1e-12, as for the phase points of the batched path.
    Membership is the oracle's FP64 cap test (cpu/grid.hpp:144-148) on its
    HEALPix pixel centres.  Caps (PPM_MEMBER_CAP) with GL or uniform
    sampling; every patch counts, visible or not."""
    import math
    L = lib()
    ns, npp = grid.n_side, grid.n_phase
    if grid.sampling == abi.SAMPLE_GL:
        pts = [L.or_gl_node(grid.n_gl, k, b / npp, (b + 1) / npp) for b in range(npp) for k in range(grid.n_gl)]
    else:
        pts = [b / npp for b in range(npp)]
    z, n_pix, s = c_double(), ctypes.c_int(), c_double()
    total = 0
    for q in range(star.n_spots):
        sp = star.spot[q]
        if sp.membership != abi.MEMBER_CAP:
            raise ValueError("caps only")
        for i in range(1, 4 * ns):
            L.or_healpix_ring(ns, i, byref(z), byref(n_pix), byref(s))
            n, zz = n_pix.value, z.value
            s2 = math.sqrt(1.0 - zz * zz)
            active = []
            for j in range(1, n + 1):
                phi = L.or_healpix_phi(ns, i, j)
                inside = math.sin(sp.theta1) * s2 * math.cos(phi - sp.phi1) + math.cos(sp.theta1) * zz > math.cos(sp.angrad1)
                if inside and sp.angrad2 > 0:
                    inside = not (math.sin(sp.theta2) * s2 * math.cos(phi - sp.phi2) + math.cos(sp.theta2) * zz
                                  > math.cos(sp.angrad2))
                if inside:
                    active.append(j)
            if not active:
                continue
            offs = []
            for p in pts:
                v = p * n
                f = math.floor(v)
                d = v - f
                if d > 1.0 - 1e-12:
                    d, f = 0.0, f + 1
                offs.append((d, int(f)))
            # classes: offsets within 1e-9 of the class's first, in ascending order
            cls, first = {}, None
            for d in sorted(set(o[0] for o in offs)):
                if first is None or d - first > 1e-9:
                    first = d
                cls[d] = first
            act = np.array(active)
            seen = set()
            for d, f in offs:
                seen.update((cls[d], int(k)) for k in (f + act) % n)
            total += len(seen)
    return total
