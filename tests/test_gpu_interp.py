"""GPU parity of the NSX interpolants (ppm_star.nsx_interp: Interp_IT3_2l2c_le,
_4c with and without use_sg, _4c_le, _4l of cpu/nsx.hpp, as t_interp selects
them, cpu/t_interp.cpp:210-219) against the FP64 oracle, on the plain and the
batched paths, in mixed batches and through the serving loop, and on the
nsx_H_v171019 axes.  Tolerance as for every profile: max relative error per
bin <= 1e-4 over bins above 1e-6 of the peak.

On the synthetic tables the interpolant classes differ by 5e-4 .. 7e-2 (so
each parity check discriminates its own mode), except 4c against 4c_le, which
differ only in the last mu brackets: a table with strong mu structure
(2 + sin 40 mu) separates them (test_mu_edge_interpolants)."""
import copy
import math
import os

import numpy as np
import pytest

from gpu_ppm_amd import abi, fixtures
from gpu_ppm_amd import scenarios as S
from gpu_ppm_amd.lib import Context, PpmError
from oracle_binding import Oracle, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4
THREADS = 16
MODES = [abi.NSX_2L2C_LE, abi.NSX_4C, abi.NSX_4C_SG, abi.NSX_4C_LE, abi.NSX_4L]
NAMES = {abi.NSX_2L2C_LE: "2l2c_le", abi.NSX_4C: "4c", abi.NSX_4C_SG: "4c_sg", abi.NSX_4C_LE: "4c_le",
         abi.NSX_4L: "4l"}
TI_ENERGIES = S.tinterp_energies()[::4]  # 256 of t_interp's 1024 energies


@pytest.fixture(scope="module")
def ctx(data_root, small_nsx_cfg):
    c = Context(data_root, nsx_path=fixtures.nsx_path(data_root, small=True), responses=("std", "std128"),
                nsx_config=small_nsx_cfg)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc(data_root, small_nsx_cfg):
    o = Oracle(data_root, nsx_path=fixtures.nsx_path(data_root, small=True), responses=("std", "std128"),
               nsx_cfg=small_nsx_cfg)
    yield o
    o.close()


def with_mode(star, mode):
    s = copy.copy(star)
    s.nsx_interp = mode
    return s


def tinterp_grid(method):
    return S.tinterp_grid(energies=TI_ENERGIES, method=method)


@pytest.mark.parametrize("mode", MODES, ids=[NAMES[m] for m in MODES])
def test_tinterp_plain(ctx, orc, mode):
    """t_interp's element (logT 6, 1024 fine phases, 256 bins x 2 GL nodes)
    on the plain path."""
    for name in ("test10", "test20"):
        st = S.tinterp_star(name, nsx_interp=mode)
        g = tinterp_grid(abi.METHOD_PLAIN)
        print("case: t_interp", name, NAMES[mode], "plain", flush=True)
        flux, _, _ = ctx.eval([st], g)
        ref, _ = orc.profile(st, g, threads=THREADS)
        assert rel_err(flux[0], ref) <= TOL, (name, NAMES[mode], rel_err(flux[0], ref))


def test_tinterp_batched_mixed(ctx, orc):
    """All five interpolants on both t_interp geometries in one batched call
    (one launch per energy-axis evaluation, results back in star order)."""
    stars = [S.tinterp_star(n, nsx_interp=m) for m in MODES for n in ("test10", "test20")]
    g = tinterp_grid(abi.METHOD_BATCHED)
    print("case: t_interp batched, 10 stars, mixed interpolants", flush=True)
    flux, _, stats = ctx.eval(stars, g)
    for k, st in enumerate(stars):
        ref, _ = orc.profile(st, g, threads=THREADS)
        assert rel_err(flux[k], ref) <= TOL, (k, NAMES[st.nsx_interp], rel_err(flux[k], ref))
    assert stats.n_eval_done > 0


def test_use_sg_clip(ctx, orc):
    """The cubic's Wien-tail undershoot (thousands of negative entries below
    the comparison floor) shows in 4c; 4c_sg returns none."""
    g = tinterp_grid(abi.METHOD_PLAIN)
    st = S.tinterp_star("test10", nsx_interp=abi.NSX_4C)
    print("case: use_sg clip", flush=True)
    f4, _, _ = ctx.eval([st], g)
    fsg, _, _ = ctx.eval([with_mode(st, abi.NSX_4C_SG)], g)
    ref, _ = orc.profile(st, g, threads=THREADS)
    n_ref = int((ref < 0).sum())
    assert n_ref > 1000
    assert (f4[0] < 0).sum() > 0.9 * n_ref
    assert (fsg[0] >= 0).all()
    fb, _, _ = ctx.eval([with_mode(st, abi.NSX_4C_SG)], tinterp_grid(abi.METHOD_BATCHED))
    assert (fb[0] >= 0).all()


@pytest.mark.parametrize("mode", MODES, ids=[NAMES[m] for m in MODES])
def test_extended_spots(ctx, orc, mode):
    """HEALPix spots (per-ring log g, disk differences, fold to counts):
    plain for one star, batched for two."""
    g = S.devarshi_grid(48, response=1)
    a = with_mode(S.devarshi_star("Crescent_Polar"), mode)
    b = with_mode(S.cfg2_star(), mode)
    print("case: extended spots", NAMES[mode], flush=True)
    fp, cp, _ = ctx.eval([a], S.devarshi_grid(48, response=1, method=abi.METHOD_PLAIN))
    fb, cb, _ = ctx.eval([a, b], S.devarshi_grid(48, response=1, method=abi.METHOD_BATCHED))
    for k, st in enumerate((a, b)):
        ref, _ = orc.profile(st, g, threads=THREADS)
        cref = orc.fold(ref, 1, st.exposure_s / g.n_phase)
        assert rel_err(fb[k], ref) <= TOL, (k, rel_err(fb[k], ref))
        assert rel_err(cb[k], cref) <= TOL
        if k == 0:
            assert rel_err(fp[0], ref) <= TOL, rel_err(fp[0], ref)
            assert rel_err(cp[0], cref) <= TOL


@pytest.mark.parametrize("mode", MODES, ids=[NAMES[m] for m in MODES])
def test_axis_clamps(ctx, orc, mode):
    """Every axis beyond the table (cpu/hunt.hpp:160-172): logT below and
    above (kT 0.005 and 1.2 keV: logT 4.8 and 7.1 against 5.1-6.8), log g
    above (override 15.5 > 15.0), logEkT beyond both ends (0.01-100 keV).  A
    wavefront's energies then span more slice rows than a row set holds, so
    the per-energy evaluation paths run too.  Plain and batched."""
    energies = S.geomspace(0.01, 100.0, 256)
    for kT, logg in ((0.005, 0.0), (1.2, 15.5)):
        st = with_mode(S.cfg1_star(), mode)
        st.spot[0].kT_keV = kT
        st.logg_override = logg
        ref = None
        for method in (abi.METHOD_PLAIN, abi.METHOD_BATCHED):
            g = abi.make_grid(32, 16, sampling=abi.SAMPLE_GL, n_gl=2, energies=energies, method=method)
            print("case: axis clamps", NAMES[mode], "kT", kT, "method", method, flush=True)
            flux, _, _ = ctx.eval([st], g)
            if ref is None:
                ref, info = orc.profile(st, g, threads=THREADS)
                assert info.n_clamped > 0
            assert rel_err(flux[0], ref) <= TOL, (NAMES[mode], kT, method, rel_err(flux[0], ref))


# ------------------------------------------------- a table with mu structure
MU_NODES = np.cos(np.arange(67) / 66 * (0.5 * math.pi - 0.0175) * (0.35 + 0.65 * np.arange(67) / 66))


@pytest.fixture(scope="module")
def wiggly(data_root, small_nsx_cfg):
    """Small-table axes; B_nu-like in logEkT times (2 + sin 40 mu) and mild
    logT / log g factors."""
    path = os.path.join(data_root, "ext", "model_data", "nsx_wiggly_mu.out")
    cfg = small_nsx_cfg
    if not os.path.exists(path):
        T = cfg.logT_min + (cfg.logT_max - cfg.logT_min) * np.arange(cfg.n_logT) / (cfg.n_logT - 1)
        G = cfg.logg_min + (cfg.logg_max - cfg.logg_min) * np.arange(cfg.n_logg) / (cfg.n_logg - 1)
        E = cfg.logEkT_min + (cfg.logEkT_max - cfg.logEkT_min) * np.arange(cfg.n_logEkT) / (cfg.n_logEkT - 1)
        t, g, e, m = np.meshgrid(T, G, E, MU_NODES, indexing="ij")
        x = 10.0 ** e
        v = x ** 3 / np.expm1(x) * (2 + np.sin(40 * m)) * (1 + 0.1 * np.sin(3 * t)) * (1 + 0.1 * np.cos(4 * g))
        rows = np.stack([e.ravel(), m.ravel(), np.log10(v.ravel()), t.ravel(), g.ravel()], axis=1)
        tmp = path + ".tmp"
        np.savetxt(tmp, rows, fmt="%.12g")
        os.replace(tmp, path)
    c = Context(data_root, nsx_path=path, responses=("std128",), nsx_config=cfg)
    o = Oracle(data_root, nsx_path=path, responses=("std128",), nsx_cfg=cfg)
    yield c, o
    c.close()
    o.close()


def test_mu_edge_interpolants(wiggly):
    """4c and 4c_le differ in the last mu brackets only (cubic vs linear);
    on this table by > 1e-3, and the GPU follows each within 1e-4."""
    c, o = wiggly
    g = S.devarshi_grid(48, response=0)
    st = S.devarshi_star("Crescent_Polar")
    refs = {}
    for mode in (abi.NSX_4C, abi.NSX_4C_LE, abi.NSX_2L2C_LE):
        s = with_mode(st, mode)
        print("case: mu-structured table", NAMES[mode], flush=True)
        refs[mode], _ = o.profile(s, g, threads=THREADS)
        fp, _, _ = c.eval([s], S.devarshi_grid(48, response=0, method=abi.METHOD_PLAIN))
        fb, _, _ = c.eval([s, s], S.devarshi_grid(48, response=0, method=abi.METHOD_BATCHED))
        assert rel_err(fp[0], refs[mode]) <= TOL, (NAMES[mode], rel_err(fp[0], refs[mode]))
        assert rel_err(fb[1], refs[mode]) <= TOL, (NAMES[mode], rel_err(fb[1], refs[mode]))
    assert rel_err(refs[abi.NSX_4C_LE], refs[abi.NSX_4C]) > 1e-3
    assert rel_err(refs[abi.NSX_4C_LE], refs[abi.NSX_2L2C_LE]) > 1e-3


# ---------------------------------------------------- nsx_H_v171019 axes
def test_v171019_config(data_root):
    """A ctx opened on the nsx_H_v171019 axes (35 x 11 x 166 x 67, log g
    13.7-14.7, cpu/nsx.hpp:39-51) against the oracle on the same file."""
    path = fixtures.ensure_nsx_v171019(data_root)
    cfg = abi.nsx_config(**abi.NSX_H_V171019)
    print("case: v171019 open", flush=True)
    c = Context(data_root, nsx_path=path, responses=("std128",), nsx_config=cfg)
    o = Oracle(data_root, nsx_path=path, responses=("std128",), nsx_cfg=cfg)
    try:
        g = S.devarshi_grid(48, response=0)
        for mode in (abi.NSX_2L2C_LE, abi.NSX_4C):
            st = with_mode(S.devarshi_star("Ring_Eq"), mode)
            print("case: v171019", NAMES[mode], flush=True)
            flux, _, _ = c.eval([st], g)
            ref, _ = o.profile(st, g, threads=THREADS)
            assert rel_err(flux[0], ref) <= TOL, (NAMES[mode], rel_err(flux[0], ref))
    finally:
        c.close()
        o.close()


# ------------------------------------------------------------- serving
def test_serving_mixed_interpolants(ctx, orc):
    """Queued requests with different interpolants run in per-class
    launches and give their own profiles."""
    g = tinterp_grid(abi.METHOD_AUTO)
    stars = [S.tinterp_star("test20", nsx_interp=m) for m in MODES]
    print("case: serving 5 interpolants", flush=True)
    tickets = [ctx.submit(st, g) for st in stars]
    ctx.drain()
    for st, t in zip(stars, tickets):
        status, flux, _, _ = ctx.poll(t, g)
        assert status == abi.OK
        ref, _ = orc.profile(st, g, threads=THREADS)
        assert rel_err(flux, ref) <= TOL, (NAMES[st.nsx_interp], rel_err(flux, ref))


def test_bad_interpolant_rejected(ctx):
    st = S.tinterp_star("test10", nsx_interp=7)
    with pytest.raises(PpmError) as e:
        ctx.eval([st], tinterp_grid(abi.METHOD_PLAIN))
    assert e.value.status == abi.ERR_ARG
