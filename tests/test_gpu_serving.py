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
    """Requests are admitted against the budget (outputs, ring tables and
    their least accumulators); the device memory the serving lanes hold
    after a stream of requests stays within it."""
    budget = 2 << 20
    with Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                 responses=("std",), nsx_config=small_nsx_cfg, hbm_budget_bytes=budget) as c:
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
        for st in stars:  # one by one: every request fits alone
            assert wait(c, c.submit(st, g), g)[0] == abi.OK
        ss = c.serve_stats()
        print("  serving lanes hold %d bytes, %d of them segment scratch (budget %d); launches %d + %d"
              % (ss.device_bytes, ss.fixed_bytes, budget, ss.launches[0], ss.launches[1]), flush=True)
        assert ss.budget_bytes == budget and ss.charged_bytes == 0 and ss.queued == ss.running == 0
        assert 0 < ss.device_bytes - ss.fixed_bytes <= budget
        # a request whose tables alone exceed the budget's launch share is refused
        big = abi.make_grid(1024, 64, sampling=abi.SAMPLE_GL, n_gl=2, response=0)
        with pytest.raises(PpmError) as ei:
            c.submit(stars[0], big)
        assert ei.value.status == abi.ERR_NOMEM


def test_light_request_overtakes_a_heavy_launch(ctx, data_root, small_nsx_cfg):
    """A light request submitted while a heavy N_side 4096 request runs on
    the device starts in the workgroup slots the heavy launch leaves free and
    finishes first; both results equal their direct evaluations and the
    light one the oracle.  (The heavy request takes about 0.1 s; at N_side
    2048 it ends too close to the light one for the order to be a test.)"""
    heavy_g = abi.make_grid(4096, 128, sampling=abi.SAMPLE_GL, n_gl=2, response=0)
    light_g = abi.make_grid(64, 16, sampling=abi.SAMPLE_GL, n_gl=2, response=1)
    base = S.sample_priors(3, seed=51)
    heavy, warm, light = base
    heavy.spot[0].angrad1 = 1.0   # about 4e12 evaluations: far above the light-work threshold
    warm.spot[0].angrad1 = 0.6
    light.spot[0].angrad1 = 0.3
    print("case: warm-up of both grids", flush=True)
    for t, g in [(ctx.submit(warm, heavy_g), heavy_g), (ctx.submit(light, light_g), light_g)]:
        assert wait(ctx, t, g)[0] == abi.OK
    print("case: a heavy request, then a light one 20 ms later", flush=True)
    t0 = time.perf_counter()
    th = ctx.submit(heavy, heavy_g)
    time.sleep(0.02)
    t_light = time.perf_counter()
    tl = ctx.submit(light, light_g)
    st, fl, cl, lat_l = wait(ctx, tl, light_g)
    assert st == abi.OK
    sh, fh_s, ch_s, lat_h = wait(ctx, th, heavy_g)
    assert sh == abi.OK
    done_light, done_heavy = t_light + lat_l * 1e-3, t0 + lat_h * 1e-3
    print("  light done at %.1f ms, heavy at %.1f ms (after the heavy submit)"
          % ((done_light - t0) * 1e3, (done_heavy - t0) * 1e3), flush=True)
    assert done_light < done_heavy
    assert ctx.serve_stats().overlapped >= 1
    fd, cd, _ = ctx.eval([light], light_g)
    assert rel_err(fl, fd[0]) <= 1e-5 and rel_err(cl, cd[0]) <= 1e-5
    fh, ch, _ = ctx.eval([heavy], heavy_g)
    assert rel_err(fh_s, fh[0]) <= 1e-5 and rel_err(ch_s, ch[0]) <= 1e-5
    orc = Oracle(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                 responses=("std", "std128"), nsx_cfg=small_nsx_cfg)
    try:
        o, _ = orc.profile(light, light_g, threads=16)
        assert rel_err(fl, o) <= 1e-4
    finally:
        orc.close()
