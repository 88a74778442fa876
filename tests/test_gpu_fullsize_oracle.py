"""The benchmark configurations at their full sizes against the FP64 oracle.

The oracle's profiles are committed fixtures (tests/golden/oracle_fullsize,
written by tests/golden/make_oracle_fullsize.py from the same seeded inputs,
whose hashes are checked here); test_oracle_fullsize_fixtures.py re-runs
columns of them live on the CPU.  Each energy column of a profile is
independent, so the fixtures of cfg1-cfg3 hold a subset of the columns and,
for cfg0-cfg2, the counts of the oracle's FP64 fold of the full profile.

  cfg0  cap, blackbody, N_side 128, 32 bins x 2 GL, 128 E: every flux column
        and the counts, plain and batched
  cfg1  cap, 35 x 14 NSX grid, N_side 1024, 64 x 2, 360 E: 37 flux columns
        and the counts, plain and batched
  cfg2  ring + crescent, N_side 2048, 128 x 2, 360 E: 37 flux columns and the
        counts, plain and batched
  cfg3  8 prior-range stars (spot radius 0.02 and 1.2 rad, kT 0.1 keV among
        them) in ONE batched launch at N_side 1024, 64 x 2, 360 E: 12 flux
        columns each
  cfg4  served requests at N_side 1024 and 2048 (the cfg1 and cfg2 stars,
        submitted together with lighter requests): the same fixtures

Tolerance: 1e-4 max relative error per bin above 1e-6 of the bin peak, for
flux and counts."""
import hashlib
import os
import time

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context
from oracle_binding import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-4
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "oracle_fullsize")


def _sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


@pytest.fixture(scope="module")
def full_root(data_root):
    fixtures.ensure_nsx(data_root, small=False)
    with open(os.path.join(FIX, "INPUTS.sha256")) as f:
        for line in f:
            digest, rel = line.split()
            assert _sha(os.path.join(data_root, rel)) == digest, "generated input differs: " + rel
    return data_root


def load(name):
    return dict(np.load(os.path.join(FIX, name + ".npz")))


def stars_of(fx):
    n = len(fx["n_active"])
    raw = fx["stars"].tobytes()
    size = len(raw) // n
    return [abi.Star.from_buffer_copy(raw[k * size:(k + 1) * size]) for k in range(n)]


@pytest.fixture(scope="module")
def ctx_std(full_root):
    c = Context(full_root, lensing="std", nsx_path=fixtures.nsx_path(full_root, small=False),
                responses=("std", "std128"))
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_wide(full_root):
    c = Context(full_root, lensing="wide", nsx_path=fixtures.nsx_path(full_root, small=False),
                responses=("std", "std128"))
    yield c
    c.close()


def grid_of(fx, method):
    return abi.make_grid(int(fx["n_side"]), int(fx["n_phase"]), sampling=abi.SAMPLE_GL, n_gl=2,
                         response=int(fx["response"]), method=method)


def check(name, fx, k, flux, counts=None):
    cols = fx["cols"]
    ef = rel_err(flux[:, cols], fx["flux"][k])
    msg = "%s star %d: flux %.2e over %d columns" % (name, k, ef, len(cols))
    ec = 0.0
    if counts is not None and "counts" in fx:
        ec = rel_err(counts, fx["counts"][k])
        msg += ", counts %.2e" % ec
    print("  " + msg, flush=True)
    assert ef <= TOL and ec <= TOL, msg


@pytest.mark.parametrize("name", ["cfg0", "cfg1", "cfg2"])
@pytest.mark.parametrize("method", [abi.METHOD_PLAIN, abi.METHOD_BATCHED])
def test_single_profile_configs(ctx_std, name, method):
    fx = load(name)
    (star,) = stars_of(fx)
    g = grid_of(fx, method)
    print("case: %s %s N_side %d" % (name, "plain" if method == abi.METHOD_PLAIN else "batched", g.n_side),
          flush=True)
    flux, counts, st = ctx_std.eval([star], g)
    assert st.n_active == fx["n_active"][0]
    check(name, fx, 0, flux[0], counts[0])


def test_cfg3_one_batched_launch(ctx_wide):
    fx = load("cfg3")
    stars = stars_of(fx)
    g = grid_of(fx, abi.METHOD_BATCHED)
    print("case: cfg3, %d stars in one batched launch" % len(stars), flush=True)
    flux, counts, st = ctx_wide.eval(stars, g)
    assert st.n_active == int(fx["n_active"].sum())
    for k in range(len(stars)):
        check("cfg3", fx, k, flux[k])
    # the plain path on the two extreme spot radii
    gp = grid_of(fx, abi.METHOD_PLAIN)
    for k in (5, 6):
        fp, _, _ = ctx_wide.eval([stars[k]], gp)
        check("cfg3 plain", fx, k, fp[0])


def test_cfg4_served_full_size(ctx_std):
    """Served requests at N_side 1024 and 2048 among lighter ones: each
    result equals the oracle's, and equals its direct evaluation."""
    f1, f2 = load("cfg1"), load("cfg2")
    heavy = [(f1, stars_of(f1)[0], grid_of(f1, abi.METHOD_AUTO)), (f2, stars_of(f2)[0], grid_of(f2, abi.METHOD_AUTO))]
    light = [(s, abi.make_grid(256, 32, sampling=abi.SAMPLE_GL, n_gl=2, response=1)) for s in S.sample_priors(4, 9)]
    print("case: cfg4 served N_side 1024 + 2048 + 4 light requests", flush=True)
    tickets = [ctx_std.submit(s, g) for _, s, g in heavy] + [ctx_std.submit(s, g) for s, g in light]
    ctx_std.drain()
    res = []
    for t, g in zip(tickets, [h[2] for h in heavy] + [lg for _, lg in light]):
        st, f, q, _ = ctx_std.poll(t, g)
        assert st == abi.OK
        res.append((f, q))
    for (fx, star, g), (f, q) in zip(heavy, res[:2]):
        check("cfg4 served N_side %d" % g.n_side, fx, 0, f, q)
    for (s, g), (f, q) in zip(light, res[2:]):
        fd, qd, _ = ctx_std.eval([s], g)
        assert rel_err(f, fd[0]) <= 1e-5 and rel_err(q, qd[0]) <= 1e-5
