"""Serving loop: queued requests on different grids are packed into one
batched launch (several grids per launch; a grid the batched kernel cannot
take runs on the plain path), complete out of order, and give the same
profiles as a direct evaluation; cancel, drain and budget admission behave
as declared in include/ppm.h."""
import time

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context, PpmError
from oracle_binding import Oracle, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                responses=("std", "std128"), nsx_config=small_nsx_cfg)
    yield c
    c.close()


def wait(ctx, ticket, grid, timeout=60.0):
    t0 = time.time()
    while True:
        st, flux, cnt, lat = ctx.poll(ticket, grid)
        if st != abi.PENDING:
            return st, flux, cnt, lat
        assert time.time() - t0 < timeout, "request did not finish"
        time.sleep(0.002)


def test_mixed_grids_match_direct_evaluation(ctx, data_root, small_nsx_cfg):
    g1 = abi.make_grid(32, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=1)
    g2 = abi.make_grid(48, 8, sampling=abi.SAMPLE_GL, n_gl=2, response=0)
    stars = S.sample_priors(10, seed=31)
    grids = [g1 if k % 2 == 0 else g2 for k in range(len(stars))]
    print("case: serving 10 requests on two grids", flush=True)
    tickets = [ctx.submit(st, g) for st, g in zip(stars, grids)]
    ctx.drain()
    results = [wait(ctx, t, g) for t, g in zip(tickets, grids)]
    orc = Oracle(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                 responses=("std", "std128"), nsx_cfg=small_nsx_cfg)
    try:
        for k, (st, g) in enumerate(zip(stars, grids)):
            status, flux, cnt, lat = results[k]
            assert status == abi.OK and lat > 0
            ref_f, ref_c, _ = ctx.eval([st], g)
            assert rel_err(flux, ref_f[0]) <= 1e-5, (k, rel_err(flux, ref_f[0]))
            assert rel_err(cnt, ref_c[0]) <= 1e-5
            if k < 2:  # against the FP64 oracle as well
                o, _ = orc.profile(st, g, threads=16)
                assert rel_err(flux, o) <= 1e-4
    finally:
        orc.close()
    # tickets are released by a successful poll
    assert ctx.poll(tickets[0], g1)[0] == abi.ERR_NOT_FOUND


def test_many_grids_in_one_launch(ctx):
    """A heavy request keeps the device busy while requests on five grids
    queue; they are evaluated together (one launch for the four batchable
    grids, the plain path for descending energies) and each equals its direct
    evaluation."""
    e50 = list(np.geomspace(0.3, 8.0, 50))
    grids = [abi.make_grid(32, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=1),
             abi.make_grid(48, 8, sampling=abi.SAMPLE_GL, n_gl=2, response=0),
             abi.make_grid(64, 4, n_fine=40, sampling=abi.SAMPLE_GL, n_gl=3, energies=e50, response=-1),
             abi.make_grid(40, 12, sampling=abi.SAMPLE_POINTS, energies=e50[::-1], response=-1),
             abi.make_grid(24, 32, sampling=abi.SAMPLE_GL, n_gl=1, response=1)]
    heavy_g = abi.make_grid(256, 32, sampling=abi.SAMPLE_GL, n_gl=2, response=0)
    stars = S.sample_priors(16, seed=41)
    print("case: serving 15 requests on 5 grids behind a heavy one", flush=True)
    heavy = ctx.submit(stars[15], heavy_g)
    todo = [(st, grids[k % len(grids)]) for k, st in enumerate(stars[:15])]
    tickets = [ctx.submit(st, g) for st, g in todo]
    ctx.drain()
    assert wait(ctx, heavy, heavy_g)[0] == abi.OK
    for k, ((st, g), t) in enumerate(zip(todo, tickets)):
        status, flux, cnt, _ = wait(ctx, t, g)
        assert status == abi.OK, k
        ref_f, ref_c, _ = ctx.eval([st], g)
        assert rel_err(flux, ref_f[0]) <= 1e-5, (k, rel_err(flux, ref_f[0]))
        if g.response >= 0:
            assert rel_err(cnt, ref_c[0]) <= 1e-5, k


def test_cancel_and_unknown_ticket(ctx):
    g = abi.make_grid(32, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=1)
    stars = S.sample_priors(6, seed=4)
    print("case: serving cancel", flush=True)
    tickets = [ctx.submit(st, g) for st in stars]
    assert ctx.cancel(tickets[-1]) == abi.OK
    ctx.drain()
    st_last = ctx.poll(tickets[-1], g)[0]
    assert st_last in (abi.ERR_CANCELLED, abi.OK, abi.ERR_NOT_FOUND)
    for t in tickets[:-1]:
        assert wait(ctx, t, g)[0] == abi.OK
    assert ctx.cancel(123456789) == abi.ERR_NOT_FOUND


def test_budget_admission(data_root, small_nsx_cfg):
    with Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                 responses=("std",), nsx_config=small_nsx_cfg, hbm_budget_bytes=4 << 20) as c:
        g = abi.make_grid(32, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=0)
        stars = S.sample_priors(8, seed=2)
        print("case: serving budget", flush=True)
        admitted, busy = [], 0
        for st in stars:
            try:
                admitted.append(c.submit(st, g))
            except PpmError as e:
                assert e.status == abi.ERR_BUSY
                busy += 1
        assert busy > 0 and admitted
        c.drain()
        for t in admitted:
            assert wait(c, t, g)[0] == abi.OK
        # memory is returned when requests finish: admission works again
        t = c.submit(stars[0], g)
        assert wait(c, t, g)[0] == abi.OK
