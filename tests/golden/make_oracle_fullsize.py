"""Writes tests/golden/oracle_fullsize/: the FP64 oracle's profiles of the
benchmark configurations at their full sizes, which the GPU tests compare the
device paths with (test_gpu_fullsize_oracle.py).

    python tests/golden/make_oracle_fullsize.py [--threads 8] [cfg ...]

Inputs are the synthetic tables regenerated by gpu_ppm_amd.fixtures (seeded;
their hashes are written to INPUTS.sha256 and checked by the tests).  Each
energy column of a profile is independent (the reference loops over E_obs
innermost, cpu/devarshi24st_cpu.cpp:231-238), so a subset of columns is an
exact restatement of those columns of the full profile.

  cfg0  cap 0.25 rad at 20 deg, blackbody kT 0.35, N_side 128, 32 bins x 2 GL,
        the 128 E_obs of the 300 x 128 response: every column, and the counts
  cfg1  the same star on the 35 x 14 NSX grid, N_side 1024, 64 x 2, 360 E:
        every column computed, the counts and 37 flux columns kept
  cfg2  ring + crescent, N_side 2048, 128 x 2, 360 E: every column computed
        (about an hour on 8 threads), the counts and 37 flux columns kept
  cfg3  8 prior-drawn stars at N_side 1024, 64 x 2, 360 E (wide lensing
        table) including spot radius 0.02 and 1.2 rad and kT 0.1 keV:
        12 energy columns each
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

from gpu_ppm_amd import abi, fixtures  # noqa: E402
from gpu_ppm_amd import scenarios as S  # noqa: E402

OUT = os.path.join(HERE, "oracle_fullsize")
DATA = os.path.join(TESTS, "_data", "root")
INPUTS = ["tables/lensing/std", "tables/lensing/wide", "tables/ism_inst/std", "tables/ism_inst/std128",
          "ext/model_data/nsx_H_v200804.out"]
COLS = [0, 1, 33, 66, 100, 133, 166, 200, 233, 266, 300, 359]


def stored_columns(n_e):
    """Flux columns kept in a fixture of a full profile: all of them up to
    128 energies, else every 10th and the last (the counts hold every
    column)."""
    return list(range(n_e)) if n_e <= 128 else sorted(set(range(0, n_e, 10)) | {n_e - 1})


def cfg3_stars():
    """Five prior draws (seed 3) and three edge stars: the smallest and the
    largest spot radius of the priors and the coolest spot."""
    stars = S.sample_priors(5, seed=3)
    base = dict(mass_msun=1.6, radius_eq_km=12.5, spin_hz=450.0, inclination_rad=0.9, distance_km=abi.KPC_KM,
                exposure_s=1e6, atmosphere=abi.ATM_NSX, variant=abi.VARIANT_OS1)
    stars.append(abi.make_star([abi.make_spot(1.1, 0.02, 0.3)], **base))
    stars.append(abi.make_star([abi.make_spot(1.9, 1.2, 0.35)], **base))
    stars.append(abi.make_star([abi.make_spot(0.6, 0.45, 0.1)], **base))
    return stars


def configs():
    """name -> (lensing, stars, grid, response index, column indices or None = all)."""
    g = lambda ns, nph, resp: abi.make_grid(ns, nph, sampling=abi.SAMPLE_GL, n_gl=2, response=resp)
    return {
        "cfg0": ("std", [S.cfg0_star()], g(128, 32, 1), None),
        "cfg1": ("std", [S.cfg1_star()], g(1024, 64, 0), None),
        "cfg2": ("std", [S.cfg2_star()], g(2048, 128, 0), None),
        "cfg3": ("wide", cfg3_stars(), g(1024, 64, 0), COLS),
    }


def ensure_inputs(root=DATA):
    fixtures.ensure_all(root, small_nsx=True, full_nsx=True)
    return root


def sha256_dir(path):
    out = []
    if os.path.isdir(path):
        for f in sorted(os.listdir(path)):
            out += sha256_dir(os.path.join(path, f))
        return out
    h = hashlib.sha256()
    with open(path, "rb") as fh:
        for chunk in iter(lambda: fh.read(1 << 20), b""):
            h.update(chunk)
    return [(h.hexdigest(), path)]


def input_hashes(root):
    lines = []
    for rel in INPUTS:
        for d, p in sha256_dir(os.path.join(root, rel)):
            lines.append("%s  %s" % (d, os.path.relpath(p, root)))
    return lines


def star_bytes(stars):
    return np.frombuffer(b"".join(bytes(s) for s in stars), dtype=np.uint8)


def main():
    from oracle_binding import Oracle
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("cfgs", nargs="*")
    a = ap.parse_args()
    root = ensure_inputs()
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, "INPUTS.sha256"), "w") as f:
        f.write("\n".join(input_hashes(root)) + "\n")
    nsx = fixtures.nsx_path(root, small=False)
    for name, (lensing, stars, grid, cols) in configs().items():
        if a.cfgs and name not in a.cfgs:
            continue
        orc = Oracle(root, lensing=lensing, nsx_path=nsx, responses=("std", "std128"))
        e_all, _ = orc.response(grid.response)
        idx = list(range(len(e_all))) if cols is None else cols
        g = abi.make_grid(grid.n_side, grid.n_phase, sampling=abi.SAMPLE_GL, n_gl=2,
                          energies=[float(e_all[i]) for i in idx])
        flux, counts, n_act = [], [], []
        t0 = time.time()
        for k, st in enumerate(stars):
            f, info = orc.profile(st, g, threads=a.threads)
            flux.append(f)
            n_act.append(info.n_active)
            if cols is None:
                counts.append(orc.fold(f, grid.response, st.exposure_s / grid.n_phase))
            print("%s star %d: n_active %d, %.1f s" % (name, k, info.n_active, time.time() - t0), flush=True)
        orc.close()
        flux = np.array(flux)
        if cols is None:
            keep = stored_columns(len(idx))
            flux, idx = flux[:, :, keep], [idx[k] for k in keep]
        out = dict(flux=flux, cols=np.array(idx), n_active=np.array(n_act), stars=star_bytes(stars),
                   n_side=grid.n_side, n_phase=grid.n_phase, response=grid.response)
        if counts:
            out["counts"] = np.array(counts)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)


if __name__ == "__main__":
    main()
