"""The oracle and the product's host code against the reference.

Fixtures (tests/golden/, written by tests/golden/make_golden.py from routines
compiled from the reference's own sources) pin: the sector lattice and ghost
fill (201/203), the radial meshes (202), the MHD pointwise routines
(660/661/662) and the 2-D EMF Riemann solvers (630/632/633).  quadpoints.json
holds the reference's committed output _T/_200mesh/test_quadpoints.out.
When the reference build itself is present (this container) the checks are
also run live at other sizes.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
VEC = np.load(os.path.join(GOLD, "ref_vectors.npz"))
P = orc.P


def _ulps(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return np.max(np.abs(a - b) / np.where(sp > 0, sp, 1.0))


@pytest.mark.parametrize("N,g", [(8, 4), (12, 3), (17, 3)])
def test_lattice_bit_exact_product_and_oracle(N, g):
    ref = VEC[f"lattice_N{N}_g{g}"]
    nf = N + 2 * g + 1
    o = np.zeros(nf * nf * 18)
    orc.olib().or_unit_lattice(N, g, P(o))
    assert np.array_equal(o, ref)
    p = cs.unit_lattice(cs.make_config(N, 1, g))
    m = np.zeros((nf, nf), bool)       # the reference leaves the corner blocks unset
    m[g:g + N + 1, :] = True
    m[:, g:g + N + 1] = True
    pr, rr = p.reshape(6, 3, nf, nf), ref.reshape(6, 3, nf, nf)
    assert np.array_equal(pr[:, :, m], rr[:, :, m])


def test_lattice_live_against_reference_build():
    lib = orc.rlib()
    if lib is None:
        pytest.skip("reference build not present (only in the build container)")
    for N, g in [(4, 3), (32, 3), (64, 4), (129, 5)]:
        nf = N + 2 * g + 1
        ref = np.zeros(nf * nf * 18)
        lib.ref_unit_lattice(ctypes.c_int(N), ctypes.c_int(g), P(ref))
        o = np.zeros_like(ref)
        orc.olib().or_unit_lattice(N, g, P(o))
        assert np.array_equal(o, ref), (N, g)


@pytest.mark.parametrize("key", ["radial_nr8_g3_log0", "radial_nr16_g4_log1"])
def test_radial_mesh(key):
    nr, g, r0, r1, lg = VEC[key + "_args"]
    cfg = cs.make_config(8, int(nr), int(g), radial_mesh=int(lg), rmin=r0, rmax=r1,
                         radial_bc=cs.RBC_PAD)
    rb = np.zeros(int(nr) + 2 * int(g) + 1)
    orc.olib().or_radial_mesh(ctypes.byref(cfg), P(rb))
    assert np.array_equal(rb, VEC[key])
    rp = np.zeros_like(rb)
    assert cs.lib().cs_build_radial_mesh(ctypes.byref(cfg), P(rp)) == 0
    assert np.array_equal(rp, VEC[key])


def test_prim_to_cons_vectors():
    gam, rhof, prsf = 5.0 / 3.0, 1e-12, 1e-9
    Vin, ent = VEC["p2c_V_in"], VEC["p2c_ent"]
    for i in range(len(Vin)):
        v = Vin[i].copy()
        u = np.zeros(8)
        orc.olib().or_prim_to_cons(gam, rhof, prsf, ent[i], P(v), P(u))
        assert _ulps(u, VEC["p2c_U"][i]) <= 1 and _ulps(v, VEC["p2c_V_out"][i]) <= 1, i


def test_cons_to_prim_vectors():
    gam, rhof, prsf = 5.0 / 3.0, 1e-12, 1e-9
    for i in range(len(VEC["c2p_U_in"])):
        u = VEC["c2p_U_in"][i].copy()
        v = np.zeros(8)
        e = np.array([VEC["c2p_ent_in"][i]])
        f = orc.olib().or_cons_to_prim(gam, rhof, prsf, P(e), P(u), P(v))
        assert f == VEC["c2p_flag"][i]
        assert _ulps(v, VEC["c2p_V"][i]) <= 1, i
        assert _ulps(u, VEC["c2p_U_out"][i]) <= 1
        assert _ulps(e, VEC["c2p_ent_out"][i]) <= 1


def test_wave_speeds_vectors():
    gam = 5.0 / 3.0
    for i, v in enumerate(VEC["wav_V"]):
        for d in range(3):
            # the x-direction routine applied to the state rotated so that d is x
            w = v.copy()
            w[[1, 1 + d]] = w[[1 + d, 1]]
            w[[5, 5 + d]] = w[[5 + d, 5]]
            a, b = np.zeros(1), np.zeros(1)
            orc.olib().or_wavspd_x(gam, P(w), P(a), P(b))
            if d == 0:
                assert _ulps([a[0], b[0]], VEC["wav_S"][i, d]) <= 1
            else:   # the swap reorders the B^2 sum, so only rounding-level agreement
                np.testing.assert_allclose([a[0], b[0]], VEC["wav_S"][i, d], rtol=1e-13, atol=1e-13)


def test_emf4_hll_vectors():
    S, th = VEC["emf4_S"], VEC["emf4_theta"]
    for i in range(len(S)):
        e = orc.olib().or_emf4_hll(P(S[i].copy()), np.cos(th[i]), np.sin(th[i]), P(VEC["emf4_ez"][i].copy()),
                                   P(VEC["emf4_bn"][i].copy()), P(VEC["emf4_by"][i].copy()))
        assert _ulps(e, VEC["emf4_hll_out"][i, 2]) <= 1, i


def test_emf4_symmetric_speeds_equal_llf_closed_form():
    """With S_R = -S_L = S_U = -S_D and equal corner E (the reference's
    USE_2D_LLF_RS_FOR_EZ path) the 4-state HLL solver reduces to the LLF
    closed form of 632_2d_rs_ez_two_d_rs_ez_4state_llf.f:75-85."""
    S, th = VEC["emf4_S"], VEC["emf4_theta"]
    for i in range(len(S)):
        m = np.abs(S[i]).max()
        s4 = np.array([m, -m, m, -m])
        ez = np.full(4, VEC["emf4_ez"][i].mean())
        e = orc.olib().or_emf4_hll(P(s4), np.cos(th[i]), np.sin(th[i]), P(ez), P(VEC["emf4_bn"][i].copy()),
                                   P(VEC["emf4_by"][i].copy()))
        bnU, bnD = VEC["emf4_bn"][i]
        byR, byL = VEC["emf4_by"][i]
        llf = ez[0] + 0.5 * m / np.sin(th[i]) * (bnD - bnU + byR - byL)
        assert abs(e - llf) <= 1e-12 * (1 + abs(llf))


def test_emf3_llf_vectors():
    for i in range(len(VEC["emf3_S6"])):
        S = np.abs(VEC["emf3_S6"][i]).max()
        e = orc.olib().or_emf3_llf(S, P(VEC["emf3_ez"][i].copy()), P(VEC["emf3_bn"][i].copy()))
        assert _ulps(e, VEC["emf3_out"][i, 2]) <= 1


def test_quadpoints_golden():
    """_T/_200mesh/test_quadpoints.out: sector 1, one zone, radii 10 and 11.
    Vertices (4 d.p.), and the B points that lie on the exact geometry: face
    centres and edge midpoints of the x/y faces, arc midpoints of the z faces.
    The z-face centre points (8.3686, 9.2055) belong to the reference's
    quadratic isoparametric surfaces, not to this build's exact metric."""
    d = json.load(open(os.path.join(GOLD, "quadpoints.json")))
    lat = np.zeros(2 * 2 * 18)
    orc.olib().or_unit_lattice(1, 0, P(lat))
    u = lat.reshape(6, 3, 2, 2)[0]             # sector 1: u[:, j, i]
    for ix, iy, iz, x, y, z in d["vert"]:
        r = 10.0 if iz == 0 else 11.0
        np.testing.assert_allclose(r * u[:, iy, ix], [x, y, z], atol=6e-5)

    def mid(a, b):
        m = a + b
        return m / np.linalg.norm(m)
    for face, i in (("xb_quadpts_patch_0", 0), ("xb_quadpts_patch_1", 1)):
        pts = np.array(d[face])
        a, b = u[:, 0, i], u[:, 1, i]
        np.testing.assert_allclose(pts[0], 10.5 * mid(a, b), atol=6e-5)
        np.testing.assert_allclose(pts[1], 10.5 * b, atol=6e-5)
        np.testing.assert_allclose(pts[2], 10.5 * a, atol=6e-5)
        np.testing.assert_allclose(pts[3], 11.0 * mid(a, b), atol=6e-5)
        np.testing.assert_allclose(pts[4], 10.0 * mid(a, b), atol=6e-5)
    for face, j in (("yb_quadpts_patch_0", 0), ("yb_quadpts_patch_1", 1)):
        pts = np.array(d[face])
        a, b = u[:, j, 0], u[:, j, 1]
        np.testing.assert_allclose(pts[0], 10.5 * mid(a, b), atol=6e-5)
        np.testing.assert_allclose(pts[1], 11.0 * mid(a, b), atol=6e-5)
        np.testing.assert_allclose(pts[2], 10.0 * mid(a, b), atol=6e-5)
        np.testing.assert_allclose(pts[3], 10.5 * b, atol=6e-5)
        np.testing.assert_allclose(pts[4], 10.5 * a, atol=6e-5)
    for face, r in (("zb_quadpts_patch_0", 10.0), ("zb_quadpts_patch_1", 11.0)):
        pts = np.array(d[face])
        np.testing.assert_allclose(pts[1], r * mid(u[:, 0, 1], u[:, 1, 1]), atol=6e-5)
        np.testing.assert_allclose(pts[2], r * mid(u[:, 0, 0], u[:, 1, 0]), atol=6e-5)
        np.testing.assert_allclose(pts[3], r * mid(u[:, 1, 0], u[:, 1, 1]), atol=6e-5)
        np.testing.assert_allclose(pts[4], r * mid(u[:, 0, 0], u[:, 0, 1]), atol=6e-5)
