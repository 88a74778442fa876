"""Batched path (persistent work queue, phase-shift dedupe, hipGraph) against
the FP64 oracle, against the plain path, and for isolation: a star's result
is the same alone and in a mix.  Tolerance as for the plain path: 1e-4 max
relative error per bin above 1e-6 of the peak."""
import os

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context
from oracle_binding import Oracle, distinct_phases, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                responses=("std", "std128"), nsx_config=small_nsx_cfg)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc(data_root, small_nsx_cfg):
    o = Oracle(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
               responses=("std", "std128"), nsx_cfg=small_nsx_cfg)
    yield o
    o.close()


def batch_grid(n_side, n_phase, response=1, method=abi.METHOD_BATCHED, **kw):
    return abi.make_grid(n_side, n_phase, sampling=abi.SAMPLE_GL, n_gl=2, response=response, method=method, **kw)


def test_batched_matches_oracle(ctx, orc):
    stars = S.sample_priors(6, seed=11)
    g = batch_grid(32, 16)
    print("case: batched priors N 32, 6 stars", flush=True)
    flux, counts, stats = ctx.eval(stars, g)
    assert stats.n_eval_done > 0
    for k, st in enumerate(stars):
        ref, info = orc.profile(st, g, threads=16)
        cref = orc.fold(ref, 1, st.exposure_s / g.n_phase)
        assert rel_err(flux[k], ref) <= TOL, (k, rel_err(flux[k], ref))
        assert rel_err(counts[k], cref) <= TOL, (k, rel_err(counts[k], cref))


def test_batched_equals_plain_and_dedupes(ctx):
    stars = S.sample_priors(5, seed=3)
    gb = batch_grid(64, 16)
    gp = batch_grid(64, 16, method=abi.METHOD_PLAIN)
    print("case: batched vs plain N 64, 5 stars", flush=True)
    fb, cb, sb = ctx.eval(stars, gb)
    fp, cp, sp = ctx.eval(stars, gp)
    # both are FP32 evaluations of the same sums (FP64 prefix sums here,
    # FP32 per-64-patch partial sums with compensated bin sums there)
    for k in range(len(stars)):
        assert rel_err(fb[k], fp[k]) <= 1e-5
        assert rel_err(cb[k], cp[k]) <= 1e-5
    assert sb.n_active == sp.n_active and sb.n_eval_ref == sp.n_eval_ref
    assert sb.n_eval_done < sb.n_eval_ref  # phase-shift dedupe and zero skips


def test_alone_and_in_mix(ctx):
    stars = S.sample_priors(7, seed=21)
    g = batch_grid(48, 16)
    print("case: alone vs mix", flush=True)
    mix, cmix, _ = ctx.eval(stars, g)
    for k in (0, 3, 6):
        alone, calone, _ = ctx.eval([stars[k]], g)
        assert rel_err(alone[0], mix[k]) <= 1e-12
        assert rel_err(calone[0], cmix[k]) <= 1e-12


def test_graph_replay_matches_direct_launch(ctx, data_root, small_nsx_cfg, monkeypatch):
    stars = S.sample_priors(4, seed=5)
    g = batch_grid(32, 16)
    f1, c1, _ = ctx.eval(stars, g)
    f2, c2, _ = ctx.eval(stars, g)   # replay of the captured graph
    f3, c3, _ = ctx.eval(stars, g)
    assert np.array_equal(f1, f2) and np.array_equal(f1, f3) and np.array_equal(c1, c3)
    monkeypatch.setenv("PPM_NO_GRAPH", "1")
    with Context(data_root, lensing="wide", nsx_path=fixtures.nsx_path(data_root, small=True),
                 responses=("std", "std128"), nsx_config=small_nsx_cfg) as c2x:
        fd, cd, _ = c2x.eval(stars, g)
    assert np.array_equal(f1, fd) and np.array_equal(c1, cd)


def test_batched_energy_chunks_and_blackbody(ctx, orc):
    """More energies than one workgroup covers (384), blackbody atmosphere,
    two-spot stars and a point source in one batch."""
    energies = S.geomspace(0.2, 12.0, 500)
    g = abi.make_grid(32, 8, sampling=abi.SAMPLE_GL, n_gl=3, energies=energies, method=abi.METHOD_BATCHED)
    stars = S.sample_priors(3, seed=8, atmosphere=abi.ATM_BLACKBODY)
    two = abi.make_star([abi.make_spot(0.7, 0.3, 0.3), abi.make_spot(2.0, 0.5, 0.2, theta2=2.1, angrad2=0.2,
                                                                    membership=abi.MEMBER_DISKDIFF)])
    pt = abi.make_star([abi.make_spot(1.2, 0.05, 0.35)], point_source_below=0.1)
    stars += [two, pt]
    print("case: 500 E blackbody batch", flush=True)
    flux, _, _ = ctx.eval(stars, g)
    for k, st in enumerate(stars):
        ref, _ = orc.profile(st, g, threads=16)
        assert rel_err(flux[k], ref) <= TOL, (k, rel_err(flux[k], ref))


def test_batched_equals_plain_full_resolution(ctx):
    """At N_side 1024 (rings of up to 4096 pixels, two phase-shift classes on
    the equatorial rings, segments split between wavefronts) the batched path
    agrees with the plain path on prior-drawn stars."""
    stars = S.sample_priors(4, seed=0)
    gb = batch_grid(1024, 64, response=0)
    gp = batch_grid(1024, 64, response=0, method=abi.METHOD_PLAIN)
    print("case: batched vs plain N 1024, 4 stars", flush=True)
    fb, cb, sb = ctx.eval(stars, gb)
    fp, cp, sp = ctx.eval(stars, gp)
    for k in range(len(stars)):
        assert rel_err(fb[k], fp[k]) <= 1e-5, (k, rel_err(fb[k], fp[k]))
        assert rel_err(cb[k], cp[k]) <= 1e-5, (k, rel_err(cb[k], cp[k]))
    assert sb.n_active == sp.n_active and sb.n_eval_ref == sp.n_eval_ref
    assert 0 < sb.n_eval_done < sb.n_eval_ref


def test_energy_chunks_after_a_larger_batch(ctx, orc):
    """Items of one energy chunk hold accumulators for their own energies
    only: a batch with more than 384 energies, run right after a larger,
    different batch on the same context (whose accumulator memory it reuses),
    still matches the oracle in every chunk."""
    first = S.sample_priors(12, seed=31)
    print("case: larger batch first (12 stars, N 64, 360 E)", flush=True)
    ctx.eval(first, batch_grid(64, 32, response=0))
    energies = S.geomspace(0.2, 12.0, 500)
    g = abi.make_grid(48, 16, sampling=abi.SAMPLE_GL, n_gl=2, energies=energies, method=abi.METHOD_BATCHED)
    stars = S.sample_priors(3, seed=32)
    print("case: then 3 stars, N 48, 500 E (two energy chunks)", flush=True)
    flux, _, _ = ctx.eval(stars, g)
    for k, st in enumerate(stars):
        ref, _ = orc.profile(st, g, threads=16)
        e = rel_err(flux[k], ref)
        print("  star %d: %.2e (high chunk %.2e)" % (k, e, rel_err(flux[k][:, 384:], ref[:, 384:])), flush=True)
        assert e <= TOL, (k, e)


def test_auto_method_for_one_star(ctx):
    """PPM_METHOD_AUTO with one star (360 energies) takes the batched path
    from N_side 128 (ascending energies), the plain path below it and for
    unordered energies: the results equal those of the method named
    explicitly."""
    (star,) = S.sample_priors(1, seed=4)
    cases = [(256, dict(response=0), abi.METHOD_BATCHED), (64, dict(response=0), abi.METHOD_PLAIN),
             (256, dict(response=-1, energies=[2.0, 1.0, 3.0]), abi.METHOD_PLAIN)]
    for n_side, kw, expect in cases:
        print("case: AUTO, one star, N_side %d, %s" % (n_side, kw), flush=True)
        fa, _, sa = ctx.eval([star], batch_grid(n_side, 16, method=abi.METHOD_AUTO, **kw))
        fe, _, se = ctx.eval([star], batch_grid(n_side, 16, method=expect, **kw))
        assert np.array_equal(fa, fe)
        assert sa.n_eval_done == se.n_eval_done
        assert (sa.n_eval_done < sa.n_eval_ref) == (expect == abi.METHOD_BATCHED)


def test_batched_empty_and_hidden_spots(ctx, orc):
    """One batched launch holding a spot smaller than a pixel (no active
    patch), a spot the observer never sees (at the south pole, observer 3 deg
    from the north pole) and an ordinary star: each equals the oracle (the
    first two are zero), and a launch of empty spots alone returns zeros."""
    g = batch_grid(32, 16)
    kw = dict(atmosphere=abi.ATM_NSX, variant=abi.VARIANT_DEVARSHI)
    empty = abi.make_star([abi.make_spot(1.0, 1e-4, 0.3)], **kw)
    hidden = abi.make_star([abi.make_spot(np.pi, 0.3, 0.3)], inclination_rad=0.05, **kw)
    (normal,) = S.sample_priors(1, seed=8)
    stars = [empty, hidden, normal]
    print("case: batched empty + hidden + ordinary spot, N 32", flush=True)
    flux, counts, stats = ctx.eval(stars, g)
    assert np.isfinite(flux).all() and np.isfinite(counts).all()
    for k, st in enumerate(stars):
        ref, info = orc.profile(st, g, threads=16)
        if k < 2:
            assert (info.n_active > 0) == (k == 1)  # the hidden spot has patches, none seen
            assert np.all(ref == 0) and np.all(flux[k] == 0) and np.all(counts[k] == 0), k
        else:
            assert rel_err(flux[k], ref) <= TOL
    assert stats.n_active > 0
    f0, c0, s0 = ctx.eval([empty, empty], g)
    assert s0.n_active == 0 and np.all(f0 == 0) and np.all(c0 == 0)


def test_256_energies_four_per_lane_matches_oracle(ctx, orc):
    """A 256-energy grid over 0.2-12 keV runs 4 energies per lane by default
    (two row sets of 128 energies within the table evaluation's slice rows):
    against the oracle, and equal to 2 per lane within 1e-5."""
    stars = S.sample_priors(4, seed=12)
    en = list(S.geomspace(0.2, 12.0, 256))
    g = abi.make_grid(32, 16, sampling=abi.SAMPLE_GL, n_gl=2, energies=en, method=abi.METHOD_BATCHED)
    print("case: batched 256 energies, N 32, 4 stars", flush=True)
    assert ctx.launch_shape(g, len(stars), abi.ATM_NSX) == (0, 0, 0)
    flux, _, _ = ctx.eval(stars, g)
    for k, st in enumerate(stars):
        ref, _ = orc.profile(st, g, threads=16)
        assert rel_err(flux[k], ref) <= TOL, (k, rel_err(flux[k], ref))
    ctx.set_launch_shape(g, len(stars), abi.ATM_NSX, 0, 0, 2)
    try:
        f2, _, _ = ctx.eval(stars, g)
    finally:
        ctx.set_launch_shape(g, len(stars), abi.ATM_NSX, 0, 0, 0)
    for k in range(len(stars)):
        assert rel_err(f2[k], flux[k]) <= 1e-5


def test_dedupe_evaluates_each_distinct_phase_once(ctx):
    """Seen from the spin axis a northern cap is visible at every phase, so
    the batched path evaluates each distinct rotational phase of the
    reference's (phase point, patch) pairs exactly once per energy:
    n_eval_done = distinct phases x energies (oracle_binding.distinct_phases).
    One star alone (its rings split by class over several items) and three
    in one launch."""
    cases = [(64, 16, (0.6, 0.4, 0.35)), (64, 32, (0.2, 1.0, 0.15)), (48, 32, (0.9, 2.5, 0.6))]
    stars, want = [], []
    for n_side, n_phase, (th, ph, rho) in cases:
        st = abi.make_star([abi.make_spot(th, rho, 0.3, phi1=ph)], inclination_rad=0.0)
        g = batch_grid(n_side, n_phase)
        n_e = ctx.n_energy(g)
        expect = distinct_phases(st, g) * n_e
        print("case: pole-on cap %s, N_side %d, %d bins: %d evaluations expected" % ((th, ph, rho), n_side, n_phase,
                                                                                    expect), flush=True)
        _, _, s = ctx.eval([st], g, counts=False)
        assert s.n_eval_done == expect
        if n_side == 64 and n_phase == 32:
            stars.append(st)
            want.append(expect)
    g = batch_grid(64, 32)
    mix = stars + [abi.make_star([abi.make_spot(0.4, 0.3, 0.25, phi1=5.0)], inclination_rad=0.0),
                   abi.make_star([abi.make_spot(1.0, 0.2, 0.35, phi1=3.0)], inclination_rad=0.0)]
    expect = sum(distinct_phases(st, g) for st in mix) * ctx.n_energy(g)
    _, _, s = ctx.eval(mix, g, counts=False)
    assert s.n_eval_done == expect
