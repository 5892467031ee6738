"""ctypes wrapper of the CPU oracle (liboracle.so) and, when it was built, of
the reference routines compiled from the reference's sources (_ref/libcsref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libcsref.so")

_o = None
_r = None


def build():
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, capture_output=True)


def olib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, dbl, i32 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.or_grid_new.restype = vp
        L.or_grid_new.argtypes = [vp]
        L.or_grid_free.argtypes = [vp]
        L.or_grid_dims.argtypes = [vp, vp]
        L.or_map_failures.argtypes = [vp]
        L.or_unit_lattice.argtypes = [i32, i32, vp]
        L.or_radial_mesh.argtypes = [vp, vp]
        L.or_metric.argtypes = [vp, vp, vp, vp, vp]
        L.or_halo.argtypes = [vp] + [vp] * 5
        L.or_fluxes.argtypes = [vp] + [vp] * 5 + [vp, vp]
        L.or_emf.argtypes = [vp] + [vp] * 5 + [vp, vp]
        L.or_cfl.argtypes = [vp, vp, vp, dbl]
        L.or_cfl.restype = dbl
        L.or_advance.argtypes = [vp, i32] + [vp] * 5 + [vp]
        L.or_stage.argtypes = [vp, i32, dbl, vp, vp, vp]
        L.or_cons_to_prim.argtypes = [dbl, dbl, dbl, vp, vp, vp]
        L.or_prim_to_cons.argtypes = [dbl, dbl, dbl, dbl, vp, vp]
        L.or_wavspd_x.argtypes = [dbl, vp, vp, vp]
        L.or_flux_x.argtypes = [vp, vp, vp]
        for f in (L.or_hll, L.or_hlld):
            f.argtypes = [dbl, dbl, dbl, vp, vp, dbl, dbl, vp, vp, vp, vp]
        L.or_weno5.argtypes = [dbl] * 5
        L.or_weno5.restype = dbl
        L.or_emf4_hll.argtypes = [vp, dbl, dbl, vp, vp, vp]
        L.or_emf4_hll.restype = dbl
        L.or_emf3_llf.argtypes = [dbl, vp, vp]
        L.or_emf3_llf.restype = dbl
        L.or_imex_pre.argtypes = [dbl, dbl, vp, vp]
        L.or_cleanup.argtypes = [vp] + [vp] * 5
        L.or_imex_combine.argtypes = [i32, dbl, dbl, vp, vp, vp, vp, vp]
        _o = L
    return _o


def rlib():
    """Reference routines (flang build of /root/reference sources) or None."""
    global _r
    if _r is None and os.path.exists(REFLIB):
        _r = ctypes.CDLL(REFLIB)
    return _r


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ptrs(arrs):
    t = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return ctypes.cast(t, ctypes.c_void_p), t


class Grid:
    """An oracle grid for a cubesphere_amd.Config."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.g = olib().or_grid_new(ctypes.byref(cfg))
        if not self.g:
            raise ValueError("oracle rejected the configuration")
        d = np.zeros(4, np.int64)
        olib().or_grid_dims(self.g, P(d))
        self.cpp, self.fpp, self.t2, self.nmap = (int(x) for x in d)

    def __del__(self):
        if getattr(self, "g", None):
            olib().or_grid_free(self.g)
            self.g = None

    def map_failures(self):
        return olib().or_map_failures(self.g)

    def halo(self, st):
        olib().or_halo(self.g, *[P(a) for a in st.arrays()])

    def fluxes(self, st):
        F = [np.zeros(9 * 6 * self.fpp) for _ in range(3)]
        S = [np.zeros(2 * 6 * self.fpp) for _ in range(3)]
        fp, _k1 = ptrs(F)
        sp, _k2 = ptrs(S)
        olib().or_fluxes(self.g, *[P(a) for a in st.arrays()], fp, sp)
        return F, S

    def emf(self, st, S):
        E = [np.zeros(6 * self.fpp) for _ in range(3)]
        sp, _k1 = ptrs(S)
        ep, _k2 = ptrs(E)
        olib().or_emf(self.g, *[P(a) for a in st.arrays()], sp, ep)
        return E

    def cfl(self, st, dt_prev):
        return olib().or_cfl(self.g, P(st.U), P(st.ent), dt_prev)

    def advance(self, st, nsteps, dt):
        d = np.array([dt])
        olib().or_advance(self.g, nsteps, *[P(a) for a in st.arrays()], P(d))
        return float(d[0])

    def stage(self, irk, dt, st_in, st_u0, st_out):
        a, _k1 = ptrs(st_in.arrays())
        b, _k2 = ptrs(st_u0.arrays())
        c, _k3 = ptrs(st_out.arrays())
        olib().or_stage(self.g, irk, dt, a, b, c)

    def cleanup(self, st):
        olib().or_cleanup(self.g, *[P(a) for a in st.arrays()])

    def metric(self):
        vol = np.zeros(6 * self.cpp)
        xa, ya, za = (np.zeros(6 * self.fpp) for _ in range(3))
        olib().or_metric(self.g, P(vol), P(xa), P(ya), P(za))
        return vol, xa, ya, za


# ------------------------------------------------------------------ quadrature geometry
GEOM_FACE_KEYS = ("pts", "wts", "normal", "area", "centroid")


def geometry_dims(cfg):
    """Per-sector element counts of the quadrature-geometry arrays (see
    cs_oracle.c): vertices, zones, x-faces, y-faces, z-faces."""
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    nv, nvr, nc, nrc = N + 2 * g + 1, NR + 2 * g + 1, N + 2 * g, NR + 2 * g
    return dict(vert=nv * nv * nvr, zone=nc * nc * nrc, xf=nv * nc * nrc, yf=nc * nv * nrc, zf=nc * nc * nvr)


def geometry_arrays(cfg):
    d = geometry_dims(cfg)
    a = dict(vert=np.zeros(6 * 3 * d["vert"]), mask=np.zeros(6 * d["zone"], np.int32),
             scale=np.zeros(6 * d["zone"]), v_pts=np.zeros(6 * 24 * d["zone"]), v_wts=np.zeros(6 * 8 * d["zone"]),
             v_tan=np.zeros(6 * 72 * d["zone"]), v_vol=np.zeros(6 * d["zone"]), v_cen=np.zeros(6 * 3 * d["zone"]),
             moments=np.zeros(6 * 4 * d["zone"]))
    for f in ("xf", "yf", "zf"):
        n = d[f]
        a[f] = dict(pts=np.zeros(6 * 15 * n), wts=np.zeros(6 * 5 * n),
                    normal=np.zeros(6 * (15 if f == "zf" else 3) * n), area=np.zeros(6 * n),
                    centroid=np.zeros(6 * 3 * n))
        b = f[0] + "b"   # B-field points of the same faces
        a[b] = dict(pts=np.zeros(6 * 15 * n), wts=np.zeros(6 * 5 * n),
                    normal=np.zeros(6 * (15 if f == "zf" else 3) * n))
    return a


def quad_geometry(cfg, use_reference=False):
    """The vertex coordinates, active mask, face and volume quadrature, scale
    length and basis moments of every sector, from the oracle or (when built)
    from the reference's own routines."""
    a = geometry_arrays(cfg)
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    vp = ctypes.c_void_p
    fptr = lambda f: [P(a[f][k]) for k in GEOM_FACE_KEYS if k in a[f]]
    if use_reference:
        R = rlib()
        if R is None:
            raise RuntimeError("reference build not present")
        lat = np.zeros((N + 2 * g + 1) ** 2 * 18)
        R.ref_unit_lattice(N, g, P(lat))
        rb = np.zeros(NR + 2 * g + 1)
        R.ref_radial_mesh(NR, g, ctypes.c_double(cfg.rmin), ctypes.c_double(cfg.rmax), 1 if cfg.radial_mesh == 1 else 0,
                          P(rb))
        R.ref_vert_coords(N, NR, g, P(lat), P(rb), P(a["vert"]), P(a["mask"]))
        R.ref_quad_faces(N, NR, g, P(a["vert"]), P(a["mask"]), *fptr("xf"), *fptr("yf"), *fptr("zf"), P(a["scale"]))
        R.ref_quad_bfield(N, NR, g, P(a["vert"]), P(a["mask"]), *fptr("xb")[:3], *fptr("yb")[:3], *fptr("zb")[:3])
        R.ref_quad_volume(N, NR, g, P(a["vert"]), P(a["mask"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_tan"]),
                          P(a["v_vol"]), P(a["v_cen"]))
        R.ref_basis_moments(N, NR, g, P(a["mask"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_vol"]), P(a["v_cen"]),
                            P(a["scale"]), P(a["moments"]))
        return a
    L = olib()
    L.or_vert_coords.argtypes = [vp, vp, vp]
    L.or_quad_faces.argtypes = [vp, vp, vp, vp, vp, vp]
    L.or_quad_volume.argtypes = [vp] * 7
    L.or_basis_moments.argtypes = [vp] * 7
    L.or_vert_coords(ctypes.byref(cfg), P(a["vert"]), P(a["mask"]))
    keep = [ptrs([a[f][k] for k in GEOM_FACE_KEYS]) for f in ("xf", "yf", "zf")]
    L.or_quad_faces(ctypes.byref(cfg), P(a["vert"]), keep[0][0], keep[1][0], keep[2][0], P(a["scale"]))
    L.or_quad_bfield.argtypes = [vp] * 5
    kb = [ptrs([a[f][k] for k in GEOM_FACE_KEYS[:3]]) for f in ("xb", "yb", "zb")]
    L.or_quad_bfield(ctypes.byref(cfg), P(a["vert"]), kb[0][0], kb[1][0], kb[2][0])
    L.or_quad_volume(ctypes.byref(cfg), P(a["vert"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_tan"]), P(a["v_vol"]),
                     P(a["v_cen"]))
    L.or_basis_moments(ctypes.byref(cfg), P(a["mask"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_cen"]), P(a["scale"]),
                       P(a["moments"]))
    return a
