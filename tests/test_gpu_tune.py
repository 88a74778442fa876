"""Autotuner of the batched launch shape: the search keeps the fastest shape
it timed for the grid shape, evaluations use it, stored shapes can be set
and cleared, and every shape gives the same profiles (only the item split
and the workgroup count change, i.e. the order of FP64 sums)."""
import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context, PpmError
from oracle_binding import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                responses=("std", "std128"), nsx_config=small_nsx_cfg)
    yield c
    c.close()


def grid(n_side=128, n_phase=32):
    return abi.make_grid(n_side, n_phase, sampling=abi.SAMPLE_GL, n_gl=2, response=1, method=abi.METHOD_BATCHED)


def test_tune_keeps_the_fastest_shape_and_results_are_unchanged(ctx):
    stars = S.sample_priors(12, seed=3)
    g = grid()
    print("case: tune N 128, 12 stars", flush=True)
    f0, c0, _ = ctx.eval(stars, g)  # default shape
    r = ctx.tune(stars, g, repeats=1)
    assert r.n_tried >= 4
    assert 0 < r.best_ms <= r.default_ms
    assert r.best.items_per_slot in (2, 5, 10, 20, 40) and r.best.wg_per_cu >= 1  # default 10, halved or doubled
    assert ctx.launch_shape(g, len(stars), abi.ATM_NSX) == (r.best.items_per_slot, r.best.wg_per_cu,
                                                            r.best.energies_per_lane)
    assert r.best.energies_per_lane == 0  # 128 energies: only the default is tried
    f1, c1, _ = ctx.eval(stars, g)  # tuned shape
    for k in range(len(stars)):
        assert rel_err(f1[k], f0[k]) <= 1e-5, k
        assert rel_err(c1[k], c0[k]) <= 1e-5, k


def test_set_and_clear_launch_shape(ctx):
    stars = S.sample_priors(5, seed=4)
    g = grid(64, 16)
    ref, _, _ = ctx.eval(stars, g)
    for items, wg in ((1, 1), (4096, 0), (0, 1)):
        print("case: shape", items, wg, flush=True)
        ctx.set_launch_shape(g, len(stars), abi.ATM_NSX, items, wg)
        assert ctx.launch_shape(g, len(stars), abi.ATM_NSX) == (items, wg, 0)
        f, _, _ = ctx.eval(stars, g)
        for k in range(len(stars)):
            assert rel_err(f[k], ref[k]) <= 1e-5, (items, wg, k)
    ctx.set_launch_shape(g, len(stars), abi.ATM_NSX, 0, 0)
    assert ctx.launch_shape(g, len(stars), abi.ATM_NSX) == (0, 0, 0)
    # a shape is kept per star-count bucket: 5 stars and 6 share one, 8 does not
    ctx.set_launch_shape(g, 5, abi.ATM_NSX, 6, 0)
    assert ctx.launch_shape(g, 6, abi.ATM_NSX) == (6, 0, 0)
    assert ctx.launch_shape(g, 8, abi.ATM_NSX) == (0, 0, 0)
    ctx.set_launch_shape(g, 5, abi.ATM_NSX, 0, 0)


def test_bad_shapes_are_rejected(ctx):
    g = grid(64, 16)
    for items, wg, epl in ((-1, 0, 0), (5000, 0, 0), (0, 65, 0), (0, 0, 3)):
        with pytest.raises(PpmError) as e:
            ctx.set_launch_shape(g, 4, abi.ATM_NSX, items, wg, epl)
        assert e.value.status == abi.ERR_ARG
    with pytest.raises(PpmError) as e:
        ctx.tune(S.sample_priors(2, seed=1), g, repeats=17)
    assert e.value.status == abi.ERR_ARG


def test_energies_per_lane(ctx):
    """200 energies: 2, 4 and 6 energies per lane (two energy chunks per
    spot, or one with idle lanes) give the same profiles; the autotuner times
    the choices whose row sets fit and keeps the fastest."""
    stars = S.sample_priors(6, seed=6)
    g = abi.make_grid(96, 16, sampling=abi.SAMPLE_GL, n_gl=2, energies=S.geomspace(0.3, 10.0, 200),
                      method=abi.METHOD_BATCHED)
    ref, _, _ = ctx.eval(stars, g)
    for epl in (2, 4, 6):
        print("case: 200 energies, %d per lane" % epl, flush=True)
        ctx.set_launch_shape(g, len(stars), abi.ATM_NSX, 0, 0, epl)
        f, _, _ = ctx.eval(stars, g)
        for k in range(len(stars)):
            assert rel_err(f[k], ref[k]) <= 1e-5, (epl, k)
    ctx.set_launch_shape(g, len(stars), abi.ATM_NSX, 0, 0, 0)
    r = ctx.tune(stars, g, repeats=1)
    assert r.n_tried >= 5 and r.best.energies_per_lane in (0, 2, 4, 6)  # another choice among the shapes timed
    print("tuned: items %d, wg %d, energies per lane %d, %.2f ms (default %.2f)"
          % (r.best.items_per_slot, r.best.wg_per_cu, r.best.energies_per_lane, r.best_ms, r.default_ms), flush=True)
