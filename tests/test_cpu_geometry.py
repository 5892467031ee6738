"""The P1 geometry (EXTRACT_VERT_COORDS_PATCH, COORDS_TO_WT_QUADPTS_FACES_ISO2,
COORDS_TO_WT_QUADPTS_VOLUME_ISO2, BUILD_MODE_INDICES, INIT_BASIS_FUNCTION_COEFS)
of the C oracle, pinned against the reference's own routines (flang build,
when present) and against what the geometry fixes: the quadratic isoparametric
zones fill the spherical shell, face weights and normals are normalised, the
radial faces' centroids lie on their spheres, and second-order moments about
the centroid vanish.  The reference STOPs for nr < NGhost, so it is compared
on shells only."""
import ctypes

import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

SHELLS = {
    "shell6x3_g3": cs.make_config(6, 3, 3, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5),
    "shell8x4_g3": cs.make_config(8, 4, 3, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.2),
    "shell5x3_g2_log": cs.make_config(5, 3, 2, radial_bc=cs.RBC_PAD, radial_mesh=cs.RMESH_LOG, rmin=1.0, rmax=2.0),
}


def _rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a - b).max())


@pytest.mark.parametrize("name", list(SHELLS))
def test_geometry_matches_reference(name):
    if orc.rlib() is None:
        pytest.skip("reference build not present")
    cfg = SHELLS[name]
    o, r = orc.quad_geometry(cfg), orc.quad_geometry(cfg, use_reference=True)
    # vertices, mask, scale length and the face quadrature: bit for bit
    for k in ("vert", "mask", "scale"):
        assert np.array_equal(o[k], r[k]), k
    for f in ("xf", "yf", "zf"):
        for k in orc.GEOM_FACE_KEYS:
            assert np.array_equal(o[f][k], r[f][k]), f"{f}.{k}"
    for f in ("xb", "yb", "zb"):   # B-field points (230)
        for k in orc.GEOM_FACE_KEYS[:3]:
            assert np.array_equal(o[f][k], r[f][k]), f"{f}.{k}"
    # volume quadrature: the points bit for bit; the Jacobians through the
    # basis derivatives at the irrational Gauss points to rounding
    assert np.array_equal(o["v_pts"], r["v_pts"])
    for k in ("v_wts", "v_tan", "v_vol", "v_cen"):
        assert _rel(o[k], r[k]) < 1e-13, k
    # first moments about the centroid are rounding-level numbers of the
    # weights' rounding: compared absolutely (dimensionless, scale-length units)
    assert np.abs(o["moments"] - r["moments"]).max() < 1e-13


def test_mode_indices():
    """Second order against the reference build; higher orders: every
    monomial of degree < order once, by degree (204build_mode_indices.f:40-150)."""
    assert cs.mode_indices(2) == [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    R = orc.rlib()
    if R is not None:
        ii, jj, kk = (np.zeros(4, np.int32) for _ in range(3))
        R.ref_mode_indices(4, orc.P(ii), orc.P(jj), orc.P(kk))
        assert list(zip(ii.tolist(), jj.tolist(), kk.tolist())) == cs.mode_indices(2)
    for order, n in ((1, 1), (2, 4), (3, 10), (4, 20)):
        m = cs.mode_indices(order)
        assert len(m) == n and len(set(m)) == n
        assert all(sum(p) < order for p in m)
        assert [sum(p) for p in m] == sorted(sum(p) for p in m)


@pytest.mark.parametrize("name", list(SHELLS) + ["surface8"])
def test_geometry_invariants(name):
    cfg = SHELLS.get(name) or cs.make_config(8, 1, 3)
    a = orc.quad_geometry(cfg)
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    nc, nrc, nv, nvr = N + 2 * g, NR + 2 * g, N + 2 * g + 1, NR + 2 * g + 1
    rb = np.zeros(nvr)
    orc.olib().or_radial_mesh(ctypes.byref(cfg), orc.P(rb))
    r0, r1 = rb[g], rb[g + NR]
    vol = a["v_vol"].reshape(6, nrc, nc, nc)[:, g:g + NR, g:g + N, g:g + N]
    exact = 4.0 * np.pi / 3.0 * (r1 ** 3 - r0 ** 3)
    assert abs(vol.sum() - exact) / exact < 0.1 * (np.pi / (2 * N)) ** 4
    for f in ("xf", "yf", "zf"):
        S = a[f]["area"].size // 6
        w = a[f]["wts"].reshape(6, 5, S)
        live = a[f]["area"].reshape(6, S) > 1e-11          # faces with four distinct vertices
        assert np.allclose(w.sum(axis=1)[live], 1.0, rtol=0, atol=1e-14)
        nrm = a[f]["normal"].reshape(6, 5 if f == "zf" else 1, 3, S)
        ln = np.sqrt((nrm ** 2).sum(axis=2))
        # unit normals on the faces of the sector's own columns (the
        # tri-corner ghost faces collapse and keep shorter or zero normals)
        assert (ln <= 1.0 + 1e-14).all()
        if f == "xf":
            own = ln.reshape(6, 1, nrc, nc, nv)[:, :, g:g + NR, g:g + N, g:g + N + 1]
        elif f == "yf":
            own = ln.reshape(6, 1, nrc, nv, nc)[:, :, g:g + NR, g:g + N + 1, g:g + N]
        else:
            own = ln.reshape(6, 5, nvr, nc, nc)[:, :, g:g + NR + 1, g:g + N, g:g + N]
        assert np.allclose(own, 1.0, rtol=0, atol=1e-14)
    # z-face centroids on their spheres
    zc = a["zf"]["centroid"].reshape(6, 3, nvr, nc, nc)
    rad = np.sqrt((zc ** 2).sum(axis=1))[:, g:g + NR + 1, g:g + N, g:g + N]
    assert np.allclose(rad, rb[g:g + NR + 1][None, :, None, None], rtol=1e-14, atol=0)
    # volume weights sum to one, first moments about the centroid vanish
    wv = a["v_wts"].reshape(6, 8, nrc, nc, nc)[:, :, g:g + NR, g:g + N, g:g + N]
    assert np.allclose(wv.sum(axis=1), 1.0, rtol=0, atol=1e-14)
    m = a["moments"].reshape(6, 4, -1)
    assert np.abs(m).max() < 1e-13
    assert (a["scale"].reshape(6, nrc, nc, nc)[:, g:g + NR, g:g + N, g:g + N] > 0).all()


def test_shell_volume_converges_fourth_order():
    """The quadratic isoparametric zones fill the shell with an error falling
    as h^4 (quadratic interpolation of the great circles)."""
    errs = []
    for N in (6, 12):
        cfg = cs.make_config(N, 3, 3, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5)
        a = orc.quad_geometry(cfg)
        g, NR, nc, nrc = 3, 3, N + 6, 9
        vol = a["v_vol"].reshape(6, nrc, nc, nc)[:, g:g + NR, g:g + N, g:g + N].sum()
        exact = 4.0 * np.pi / 3.0 * (1.5 ** 3 - 1.0)
        errs.append(abs(vol - exact) / exact)
    assert 12.0 < errs[0] / errs[1] < 20.0, errs


@pytest.mark.parametrize("N,g", [(4, 3), (8, 3), (12, 4), (16, 6)])
def test_stencil_indexing(N, g):
    """STENCIL_INDEXING_TRICORNERS: the library's table (the left-down remaps
    mirrored onto the other corners) equals the oracle's literal restatement,
    which equals the reference build; and every one-step neighbour it names
    around the tri-corners is a zone adjacent on the sphere."""
    cfg = cs.make_config(N, 3, g, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5)
    got = cs.stencil_indexing(cfg)
    L = orc.olib()
    L.or_stencil_indexing.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    o = np.zeros_like(got)
    L.or_stencil_indexing(N, g, orc.P(o))
    assert np.array_equal(got, o)
    R = orc.rlib()
    if R is not None:
        r = np.zeros_like(got)
        R.ref_stencil_indexing(N, 3, g, orc.P(r))
        assert np.array_equal(o, r)
    # adjacency on the sphere: zone centres from the vertices of sector 0
    a = orc.quad_geometry(cfg)
    nc, nv = N + 2 * g, N + 2 * g + 1
    vert = a["vert"].reshape(6, 3, 3 + 2 * g + 1, nv, nv)[0, :, g]  # (3, j, i) at k = 0
    mask = a["mask"].reshape(6, 3 + 2 * g, nc, nc)[0, g]
    cen = 0.25 * (vert[:, :-1, :-1] + vert[:, 1:, :-1] + vert[:, :-1, 1:] + vert[:, 1:, 1:])  # (3, y, x)
    st = got.reshape(12, 2, nc, nc)
    h = np.pi / 2 / N
    checked = 0
    for y in range(nc):
        for x in range(nc):
            if not mask[y, x]:
                continue
            for d in range(4):
                vx, vy = st[d, 0, y, x] - (1 - g), st[d, 1, y, x] - (1 - g)
                if (vx, vy) == (x, y) or not (0 <= vx < nc and 0 <= vy < nc) or not mask[vy, vx]:
                    continue
                dist = np.linalg.norm(cen[:, y, x] - cen[:, vy, vx])
                assert 0.4 * h < dist < 1.6 * h, (x, y, d, vx, vy, dist / h)
                checked += 1
    assert checked > 4 * N * N


# ---------------------------------------------------------------------------
# The separable form of the same geometry (cs_p1geo.h): the P1 stage of a
# shell evaluates every face's points, weights, normals, area and centroid
# from per-column tables and the radial vertices instead of storing them.
# Exact in exact arithmetic (the element map of a lateral face is r(t) u(s),
# of a radial face rb u(s, t)); compared here with the oracle's per-face
# quadrature (pinned bit for bit to the reference's 220/230 above) on every
# face of an active zone, and the zone volume, centroid and scale length
# (evaluated from the vertices as 240 does) bit for bit.

SEPARABLE = {
    "shell6x3_g4": cs.make_config(6, 3, 4, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5),
    "surface8_g4": cs.make_config(8, 1, 4),
    "shell5x3_g4_log": cs.make_config(5, 3, 4, radial_bc=cs.RBC_PAD, radial_mesh=cs.RMESH_LOG, rmin=1.0, rmax=2.0),
}


@pytest.mark.parametrize("name", list(SEPARABLE))
def test_separable_face_geometry_matches_quadrature(name):
    cfg = SEPARABLE[name]
    o = orc.quad_geometry(cfg)
    d = orc.geometry_dims(cfg)
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    nz, nf = N + 2 * g, N + 2 * g + 1
    act2 = o["mask"].reshape(6, NR + 2 * g, nz, nz)[0, 0].astype(bool)   # zone columns (j, i), same in every sector
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    for fam, f in enumerate(("xf", "yf", "zf")):
        n = d[f]
        G = dict(pts=np.zeros(6 * 15 * n), wts=np.zeros(6 * 5 * n), normal=np.zeros(6 * (15 if fam == 2 else 3) * n),
                 area=np.zeros(6 * n), centroid=np.zeros(6 * 3 * n), bw=np.zeros(6 * 5 * n))
        assert cs.lib().cs_p1s_face_geometry(ctypes.byref(cfg), fam, P(G["pts"]), P(G["wts"]), P(G["normal"]),
                                             P(G["area"]), P(G["centroid"]), P(G["bw"])) == 0
        # faces with an active zone on either side
        if fam == 0:
            a = np.zeros((nz, nf), bool)
            a[:, 1:] |= act2
            a[:, :-1] |= act2
            m = np.broadcast_to(a, (NR + 2 * g, nz, nf)).ravel()
        elif fam == 1:
            a = np.zeros((nf, nz), bool)
            a[1:, :] |= act2
            a[:-1, :] |= act2
            m = np.broadcast_to(a, (NR + 2 * g, nf, nz)).ravel()
        else:
            m = np.broadcast_to(act2, (NR + 2 * g + 1, nz, nz)).ravel()
        for k, per in (("pts", 15), ("wts", 5), ("centroid", 3), ("normal", 15 if fam == 2 else 3), ("area", 1)):
            a = G[k].reshape(6, per, n)[:, :, m]
            b = o[f][k].reshape(6, per, n)[:, :, m]
            assert _rel(a, b) < 2e-14, (f, k, _rel(a, b))
        a = G["bw"].reshape(6, 5, n)[:, :, m]
        b = o[f[0] + "b"]["wts"].reshape(6, 5, n)[:, :, m]
        assert np.abs(a - b).max() < 2e-14, (f, "bw")
    cen, vol, scl = np.zeros(6 * 3 * d["zone"]), np.zeros(6 * d["zone"]), np.zeros(6 * d["zone"])
    assert cs.lib().cs_p1s_zone_geometry(ctypes.byref(cfg), P(cen), P(vol), P(scl)) == 0
    act = o["mask"].astype(bool)
    assert np.array_equal(vol[act], o["v_vol"][act])
    assert np.array_equal(scl[act], o["scale"][act])
    c3, o3 = cen.reshape(6, 3, -1), o["v_cen"].reshape(6, 3, -1)
    am = act.reshape(6, -1)
    for s in range(6):
        assert np.array_equal(c3[s][:, am[s]], o3[s][:, am[s]])
