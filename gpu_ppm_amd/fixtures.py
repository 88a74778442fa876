"""Synthetic inputs in the reference's file layout, generated on demand by
the ppm_synth tool (gpu_ppm_amd/synth/ppm_synth.c).

Layout under a root directory (what the drop-in programs read):
  tables/lensing/std      u 0.3-0.5 x 256, 512 cos(psi) / cos(alpha) nodes
                          (the 'std' setting of c_lensing.py:8-27)
  tables/lensing/wide     u 0.15-0.55 x 256, 512 nodes (covers the batched
                          priors; u < 0.568 keeps single images)
  tables/ism_inst/std     300 channels x 360 E_obs, seed 0
  tables/ism_inst/std128  300 channels x 128 E_obs, seed 0
  ext/model_data/nsx_H_v200804.out   35x14x166x67 NSX-format grid, seed 1
  ext/model_data/nsx_small.out       6x5x166x67 grid on the same axes, seed 1
  ext/model_data/nsx_H_v171019.out   35x11x166x67 grid on the nsx_H_v171019
                                     axes (log g 13.7-14.7), seed 1
"""
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
SYNTH = os.path.join(PKG, "_bin", "ppm_synth")

SMALL_NSX_DIMS = (6, 5, 166, 67)


def _synth(*args):
    if not os.path.exists(SYNTH):
        raise RuntimeError("ppm_synth is not built; run __graft_entry__.build()")
    subprocess.run([SYNTH] + [str(a) for a in args], check=True)


def _publish(path, make):
    """Runs make(tmp) on a private temporary path and renames the result to
    path, so that processes generating the same input at once (pytest
    workers) never read a half-written file or directory."""
    if os.path.exists(path):
        return
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = "%s.tmp%d" % (path, os.getpid())
    make(tmp)
    try:
        os.rename(tmp, path)
    except OSError:  # another process published it first
        if not os.path.exists(path):
            raise
        if os.path.isdir(tmp):
            shutil.rmtree(tmp)
        else:
            os.remove(tmp)


def ensure_lensing(root, setting="std"):
    d = os.path.join(root, "tables", "lensing", setting)
    args = {"std": (0.3, 0.5, 256, 512), "wide": (0.15, 0.55, 256, 512), "coarse": (0.3, 0.5, 64, 128)}
    if setting not in args:
        raise ValueError(setting)
    _publish(d, lambda t: _synth("lensing", t, *args[setting]))
    return d


def ensure_response(root, setting="std"):
    d = os.path.join(root, "tables", "ism_inst", setting)
    sizes = {"std": (300, 360), "std128": (300, 128), "ref270": (270, 128)}[setting]
    _publish(d, lambda t: _synth("response", t, sizes[0], sizes[1], 0))
    return d


def nsx_path(root, small=False):
    name = "nsx_small.out" if small else "nsx_H_v200804.out"
    return os.path.join(root, "ext", "model_data", name)


def ensure_nsx(root, small=False):
    p = nsx_path(root, small)
    if small:
        _publish(p, lambda t: _synth("nsx", t, 1, *SMALL_NSX_DIMS))
    else:
        _publish(p, lambda t: _synth("nsx", t, 1))
    return p


def nsx_bb_path(root):
    return os.path.join(root, "ext", "model_data", "nsx_blackbody.out")


def ensure_nsx_bb(root):
    """NSX-format grid on the reduced 6x5x166x67 shape (nsx_H_v200804 axes)
    filled with the isotropic Planck intensity alone: logInuT3 =
    log10(2 k^3 x^3 / (h^2 c^2 (e^x - 1))), x = 10^logEkT."""
    p = nsx_bb_path(root)
    _publish(p, lambda t: _synth("nsx_bb", t, *SMALL_NSX_DIMS))
    return p


def ensure_nsx_v171019(root):
    """NSX-format grid on the nsx_H_v171019 axes (35x11x166x67, log g
    13.7-14.7, cpu/nsx.hpp:39-51), seed 1."""
    p = os.path.join(root, "ext", "model_data", "nsx_H_v171019.out")
    _publish(p, lambda t: _synth("nsx", t, 1, 35, 11, 166, 67, 14.7))
    return p


def small_nsx_config():
    from . import abi
    nT, ng, nE, nm = SMALL_NSX_DIMS
    return abi.nsx_config(n_logT=nT, n_logg=ng, n_logEkT=nE, n_mu=nm)


def ensure_all(root, small_nsx=True, full_nsx=False, lensing=("std", "wide"),
               responses=("std", "std128")):
    for s in lensing:
        ensure_lensing(root, s)
    for r in responses:
        ensure_response(root, r)
    if small_nsx:
        ensure_nsx(root, small=True)
    if full_nsx:
        ensure_nsx(root, small=False)
    return root
