"""The C ABI: every symbol include/ppm.h declares is exported, the ctypes
mirror has the C layouts, and host-side errors become status codes (these
paths need no GPU: tables are parsed before the device is touched)."""
import ctypes
import os
import re
import subprocess

import pytest

from gpu_ppm_amd import abi, fixtures, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ppm.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(ppm_\w+)\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    L = lib.load()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ppm_\w+)", out))
    assert set(names) <= exported


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "ppm.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ppm_spot), '
                   'sizeof(ppm_star), sizeof(ppm_grid), sizeof(ppm_nsx_config), sizeof(ppm_open_opts), '
                   'sizeof(ppm_stats), offsetof(ppm_star, spot), offsetof(ppm_grid, energies_keV), '
                   'sizeof(ppm_launch_shape), sizeof(ppm_tune_result), offsetof(ppm_tune_result, n_tried), '
                   'sizeof(ppm_serve_stats), offsetof(ppm_serve_stats, device_bytes));'
                   'return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(abi.Spot), ctypes.sizeof(abi.Star), ctypes.sizeof(abi.Grid),
            ctypes.sizeof(abi.NsxConfig), ctypes.sizeof(abi.OpenOpts), ctypes.sizeof(abi.Stats),
            abi.Star.spot.offset, abi.Grid.energies_keV.offset, ctypes.sizeof(abi.LaunchShape),
            ctypes.sizeof(abi.TuneResult), abi.TuneResult.n_tried.offset, ctypes.sizeof(abi.ServeStats),
            abi.ServeStats.device_bytes.offset]
    assert got == want


def test_strerror_covers_every_status():
    L = lib.load()
    assert L.ppm_abi_version() == 2
    for s in range(0, 11):
        assert L.ppm_strerror(s) and L.ppm_strerror(s) != b"unknown status"
    assert L.ppm_strerror(99) == b"unknown status"


def _open(root, **kw):
    L = lib.load()
    o = abi.OpenOpts()
    o.root = root.encode()
    o.lensing_setting = kw.get("lensing", "std").encode()
    o.nsx_path = kw.get("nsx", "").encode()
    o.response_settings = kw.get("responses", "").encode()
    cfg = kw.get("cfg")
    if cfg is not None:
        o.nsx_config = ctypes.pointer(cfg)
    h = ctypes.c_void_p()
    st = L.ppm_open(ctypes.byref(o), ctypes.byref(h))
    if st == abi.OK:
        L.ppm_close(h)
    return st


def test_open_missing_tables_is_io_error(tmp_path):
    assert _open(str(tmp_path)) == abi.ERR_IO
    assert _open(str(tmp_path), lensing="nope") == abi.ERR_IO


def test_open_malformed_tables(tmp_path, data_root):
    # short lensing matrix
    d = tmp_path / "tables" / "lensing" / "std"
    d.mkdir(parents=True)
    for name in ("u.txt", "cos_psi.txt", "cos_alpha.txt"):
        (d / name).write_text(open(os.path.join(data_root, "tables", "lensing", "std", name)).read())
    for name in ("cos_alpha_of_u_cos_psi.txt", "lf_of_u_cos_psi.txt", "cdt_over_R_of_u_cos_alpha.txt"):
        (d / name).write_text("1 2 3\n")
    assert _open(str(tmp_path)) == abi.ERR_FORMAT
    # bad axis
    (d / "u.txt").write_text("0.3 0.5 1\n")
    assert _open(str(tmp_path)) == abi.ERR_FORMAT


def test_open_nsx_consistency_checks(tmp_path, data_root, small_nsx_cfg):
    """The NSX reader enforces the reference loader's checks
    (cpu/nsx.hpp:110-147): rows on the uniform axes, mu consistent across
    blocks and decreasing."""
    src = fixtures.nsx_path(data_root, small=True)
    lines = open(src).read().splitlines()
    bad_axis = list(lines)
    f = bad_axis[5].split()
    f[0] = "%.10g" % (float(f[0]) + 1e-3)   # logEkT off its node
    bad_axis[5] = " ".join(f)
    p1 = tmp_path / "bad_axis.out"
    p1.write_text("\n".join(bad_axis) + "\n")
    bad_mu = list(lines)
    f = bad_mu[67 + 3].split()                # second block, mu index 3
    f[1] = "%.10g" % (float(f[1]) * 1.01)
    bad_mu[67 + 3] = " ".join(f)
    p2 = tmp_path / "bad_mu.out"
    p2.write_text("\n".join(bad_mu) + "\n")
    p3 = tmp_path / "short.out"
    p3.write_text("\n".join(lines[:1000]) + "\n")
    for p in (p1, p2, p3):
        assert _open(data_root, nsx=str(p), cfg=small_nsx_cfg) == abi.ERR_FORMAT


def test_open_without_device_reports_device_error(data_root, small_nsx_cfg):
    st = _open(data_root, nsx=fixtures.nsx_path(data_root, small=True), cfg=small_nsx_cfg,
               responses="std")
    assert st in (abi.OK, abi.ERR_DEVICE)
    if not os.path.exists("/dev/kfd"):  # no AMD GPU driver on this host
        assert st == abi.ERR_DEVICE


def test_context_raises_loudly(tmp_path):
    with pytest.raises(lib.PpmError) as e:
        lib.Context(root=str(tmp_path))
    assert e.value.status == abi.ERR_IO


def test_tuner_entry_points_reject_bad_arguments():
    """The autotuner's entry points fail with a status on a null context
    (no device needed)."""
    L = lib.load()
    g = abi.make_grid(16, 8)
    st = abi.make_star([abi.make_spot(0.5, 0.2, 0.3)])
    r = abi.TuneResult()
    sh = abi.LaunchShape(12, 0)
    assert L.ppm_tune(None, ctypes.byref(st), 1, ctypes.byref(g), 0, ctypes.byref(r)) == abi.ERR_ARG
    assert L.ppm_set_launch_shape(None, ctypes.byref(g), 1, abi.ATM_NSX, ctypes.byref(sh)) == abi.ERR_ARG
    assert L.ppm_get_launch_shape(None, ctypes.byref(g), 1, abi.ATM_NSX, ctypes.byref(sh)) == abi.ERR_ARG


def test_fit_entry_point_rejects_bad_arguments():
    L = lib.load()
    g = abi.make_grid(16, 8, response=0)
    st = abi.make_star([abi.make_spot(0.5, 0.2, 0.3)])
    fit = abi.Fit()
    data = (ctypes.c_double * 4)()
    assert L.ppm_eval_fit(None, ctypes.byref(st), 1, ctypes.byref(g), data, 0, ctypes.byref(fit), None,
                          None) == abi.ERR_ARG


def test_phase_timer_build_has_the_same_entry_points():
    """libppm_phase.so (the batched integral with phase timers) exports every
    declared symbol too, and the phase count agrees with PPM_PHASE_COUNT."""
    assert os.path.exists(lib.PHASE_LIB_PATH)
    out = subprocess.run(["nm", "-D", "--defined-only", lib.PHASE_LIB_PATH], capture_output=True, text=True).stdout
    assert set(declared_functions()) <= set(re.findall(r" T (ppm_\w+)", out))
    n = int(re.search(r"#define PPM_PHASE_COUNT (\d+)", open(HEADER).read()).group(1))
    assert n == abi.PHASE_COUNT == len(abi.PHASE_NAMES) + 1
