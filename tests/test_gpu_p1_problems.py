"""GPU parity of cs_init_problem_p1 / cs_add_boundary_p1 (INIT_PATCH_CUBE_SPHERE
820init_patch_cube_sphere.f:8 and ADD_BOUNDARY 825add_boundary.f:12 with
PROBLEM_GENERATOR_U_A 801problem_generator_u_a.f:10, on the device) against
the oracle's restatement, which is bit-exact with the reference's routines
built from source (tests/test_cpu_problems.py).

Tolerance: the device evaluates the separable form of the face areas and the
zone quadrature (a few ulps from the oracle's per-face arrays) and its own
libm (exp, pow, acos, atan2), so every array agrees with the oracle to 1e-12
of its largest magnitude; the discrete divergence of every zone is rounding
(1e-13) on the P1 geometry, also after steps."""
import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _cfg(name):
    if name == "surface":
        return cs.make_config(8, 1, 4, cs.RECON_P1ADP, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3)
    return cs.make_config(8, 3, 4, cs.RECON_P1ADP, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                          cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1)


CASES = [("shell", 0), ("shell", 2), ("shell", 103), ("surface", 102)]


def _close(got, ref, tol=1e-12):
    """U and the entropy to tol of their largest magnitude, the three face-B
    families to tol of the largest face B (a family can be zero up to
    rounding: the radial faces of the surface blast)"""
    bscale = max(max(np.abs(b).max() for b in ref[2:]), 1e-300)
    for nm, a, b in zip("U ent Bx By Bz".split(), got, ref):
        scale = bscale if nm.startswith("B") else max(np.abs(b).max(), 1e-300)
        err = np.abs(a - b).max() / scale
        assert err < tol, (nm, err)


@pytest.mark.parametrize("geom,problem", CASES, ids=[f"{g}_p{p}" for g, p in CASES])
def test_init_problem_matches_oracle(geom, problem):
    cfg = _cfg(geom)
    print(f"case {geom} problem {problem}", flush=True)
    ref = orc.init_problem_p1(cfg, problem, seed=7)
    h = cs.CubeSphere(cfg)
    try:
        h.init_problem_p1(problem, seed=7)
        got = h.download().arrays()
        divb = h.max_divb()
    finally:
        h.close()
    _close(got, ref)
    assert divb < 1e-13, divb


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_add_boundary_matches_oracle(layers):
    cfg = _cfg("shell")
    base = orc.init_problem_p1(cfg, 0)
    ref = orc.init_problem_p1(cfg, 2, layers=layers, state=base)
    h = cs.CubeSphere(cfg)
    try:
        h.init_problem_p1(0)
        h.add_boundary_p1(layers & 1, layers & 2, 2)
        got = h.download().arrays()
    finally:
        h.close()
    _close(got, ref)


def test_steps_from_the_device_initial_state_keep_div_b():
    """the hydrostatic shell (the C5 problem) from the device initial state:
    the discrete divergence stays at rounding over steps"""
    cfg = _cfg("shell")
    h = cs.CubeSphere(cfg)
    try:
        h.init_problem_p1(103, seed=3)
        d0 = h.max_divb()
        h.dt = 0.0
        h.evaluate_timestep()
        h.advance(3)
        d1 = h.max_divb()
    finally:
        h.close()
    assert d0 < 1e-13 and d1 < 1e-13, (d0, d1)


@pytest.mark.parametrize("N,NR", [(8, 3), (12, 2)])
def test_device_exchange_of_closed_form_fields(N, NR):
    """The device exchange (the one every upload and every P1 stage runs)
    against the closed forms of problem 0 built on the device: after an
    upload whose lateral ghosts are NaN, every ghost strip zone and face
    holds its own closed-form value (every component, every face-B swap and
    sign flip; tests/test_cpu_problems.py holds the comparison)."""
    from test_cpu_problems import _poison_lateral_ghosts, closed_form_exchange_check
    cfg = cs.make_config(N, NR, 4, cs.RECON_P1ADP, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                         cs.RMESH_UNIFORM, 1.0, 1.5)
    print(f"case N={N} NR={NR}", flush=True)
    h = cs.CubeSphere(cfg)
    try:
        h.init_problem_p1(0)
        full = h.download()
        st = cs.HostState(cfg)
        for a, b in zip(st.arrays(), _poison_lateral_ghosts(cfg, full.arrays())):
            a[:] = b
        h.upload(st)
        got = h.download()
    finally:
        h.close()
    dz, df, n = closed_form_exchange_check(cfg, got.arrays(), full.arrays())
    assert n > 0 and dz < 1e-14 and df < 1e-12, (dz, df)
