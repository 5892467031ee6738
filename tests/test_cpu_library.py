"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares, its host-only functions agree with the oracle, and it
reports errors the way the header documents."""
import ctypes
import os
import re

import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cubesphere_hip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^int (cs_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(cs.LIBPATH)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_codes_and_messages():
    bad = cs.make_config(16, 1, 3)
    bad.nghost = 2
    s = cs.Sizes()
    assert cs.lib().cs_sizes_of(ctypes.byref(bad), ctypes.byref(s)) == -1
    assert "nghost" in cs.last_error()
    dec = cs.make_config(5, 1, 5)          # NGhost >= ntp: FILL_GHOST2D_GLOBAL's STOP
    assert cs.lib().cs_sizes_of(ctypes.byref(dec), ctypes.byref(s)) == -2
    neg = cs.make_config(16, 1, 3)
    neg.rmax = 0.5
    assert cs.lib().cs_sizes_of(ctypes.byref(neg), ctypes.byref(s)) == -1
    assert cs.lib().cs_init_problem(ctypes.byref(cs.make_config(16, 1, 3)), 9, 0, None, None, None, None,
                                    None) == -1


def test_sizes():
    cfg = cs.make_config(16, 2, 3)
    s = cs.sizes(cfg)
    assert s.n_cell == 6 * 22 * 22 * 8 and s.n_face == 6 * 23 * 23 * 9 and s.n_interior == 6 * 16 * 16 * 2


def test_device_entry_points_fail_cleanly_without_gpu():
    import torch  # noqa: F401  (only to ask whether a device exists)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = cs.lib().cs_create(ctypes.byref(cs.make_config(8, 1, 3)), ctypes.byref(h))
    assert rc == -3 and not h.value
    assert cs.lib().cs_halo_exchange(None) == -1


@pytest.mark.parametrize("cfg", [cs.make_config(12, 1, 3), cs.make_config(10, 3, 4, radial_bc=cs.RBC_PAD)])
def test_host_metric_equals_oracle_bitwise(cfg):
    P = cs.metric(cfg)
    O = orc.Grid(cfg).metric()
    for a, b in zip(P, O):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("problem", [0, 1, 2, 3])
def test_initial_face_field_is_divergence_free(problem):
    """Face B from the vector potential by Stokes (820init_patch_cube_sphere.f:464+):
    the discrete divergence of every zone vanishes to roundoff."""
    cfg = cs.make_config(10, 2, 3, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.4)
    st = cs.init_problem(cfg, problem, seed=4)
    vol, xa, ya, za = cs.metric(cfg)
    g, N, NR = cfg.nghost, cfg.ntp, cfg.nr
    nf, nrf = N + 2 * g + 1, NR + 2 * g + 1
    Bx, By, Bz = (a.reshape(6, nrf, nf, nf) for a in st.arrays()[2:])
    Ax, Ay, Az = (a.reshape(6, nrf, nf, nf) for a in (xa, ya, za))
    i, im = slice(g + 1, g + N + 1), slice(g, g + N)
    worst = 0.0
    for k in range(1, NR + 1):
        kk = g + k
        fx = Ax[:, kk] * Bx[:, kk]
        fy = Ay[:, kk] * By[:, kk]
        div = (fx[:, i, i] - fx[:, i, im]) + (fy[:, i, i] - fy[:, im, i]) + \
              (Az[:, kk] * Bz[:, kk] - Az[:, kk - 1] * Bz[:, kk - 1])[:, i, i]
        scale = np.abs(fx[:, i, i]).max() + np.abs(fy[:, i, i]).max() + np.abs(Az[:, kk] * Bz[:, kk]).max()
        worst = max(worst, np.abs(div).max() / max(scale, 1e-300))
    assert worst < 1e-13


@pytest.mark.parametrize("N,g", [(8, 3), (9, 3), (12, 2), (16, 4), (7, 1), (32, 3)])
def test_ghost_push_maps_invert_the_exchange(N, g, tmp_path):
    """The inverse ghost maps the zone and face updates write through
    (cs_halo_maps.h) cover exactly the slots of the exchange's forward maps,
    from the same owned slot with the same sign: built from the header with
    g++ and checked exhaustively over every sector, strip, depth and family."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not installed")
    exe = tmp_path / "halo_maps_check"
    src = os.path.join(ROOT, "tests", "halo_maps_check.cpp")
    inc = os.path.join(ROOT, "cubesphere_amd", "csrc")
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", inc, src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), str(N), str(g)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


def test_third_order_entry_points_reject_null_arguments():
    """cs_limit_all_slopes_p12 and its timed form return CS_E_INVALID with a
    message, before any device call, when the patch, its moments or the
    timing array is missing."""
    L = cs.lib()
    assert L.cs_limit_all_slopes_p12(None, None) == -1
    assert "cs_limit_all_slopes_p12" in cs.last_error()
    q = cs.P12Patch()  # moments left null
    assert L.cs_limit_all_slopes_p12(ctypes.byref(q), None) == -1
    assert "moments" in cs.last_error()
    ms = (ctypes.c_float * 8)()
    assert L.cs_limit_all_slopes_p12_timed(ctypes.byref(q), None, ms) == -1
    assert L.cs_limit_all_slopes_p12_timed(None, None, None) == -1
    assert "kernel_ms" in cs.last_error()
