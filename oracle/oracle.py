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


_r_omp = None


def rlib(openmp=False):
    """Reference routines (flang build of /root/reference sources) or None;
    openmp=True: the same sources built with their OpenMP directives on."""
    global _r, _r_omp
    if openmp:
        path = os.path.join(os.path.dirname(REFLIB), "libcsref_omp.so")
        if _r_omp is None and os.path.exists(path):
            _r_omp = ctypes.CDLL(path)
        return _r_omp
    if _r is None and os.path.exists(REFLIB):
        _r = ctypes.CDLL(REFLIB)
    return _r


_r3 = None


def rlib3():
    """The reference's third-order routines (NUMORDER 3 build, libcsref3.so)
    or None."""
    global _r3
    path = os.path.join(os.path.dirname(REFLIB), "libcsref3.so")
    if _r3 is None and os.path.exists(path):
        _r3 = ctypes.CDLL(path)
    return _r3


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


# ---------------------------------------------------------------------------
# LIMIT_ALL_SLOPES_P1 on one patch (or_limit_all_slopes_p1 / the reference's
# routine / cubesphere_amd.limit_all_slopes_p1 share this structure layout).

P1_GEOM_FIELDS = ("vert", "mask", "stencil", "xf_normal", "xf_area", "xf_centroid", "yf_normal", "yf_area",
                  "yf_centroid", "zf_area", "zf_centroid", "xb_pts", "xb_wts", "xb_normal", "yb_pts", "yb_wts",
                  "yb_normal", "zb_pts", "zb_wts", "zb_normal", "scale", "v_centroid")
P1_STATE_FIELDS = ("U", "ent", "Bxf", "Byf", "Bzf", "V", "B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc",
                   "gamma_eos", "gamma_sound", "mean_mol_wt", "divvel", "msonic", "flatten", "work",
                   "Bxf_q", "Byf_q", "Bzf_q")
P1_SCALARS = ("gamma", "molewt", "rhofloor", "prsfloor", "smallnum", "flatten_coef", "prs_flatten_coef")


class P1Patch(ctypes.Structure):
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("itri", ctypes.c_int32 * 4), ("pad", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in P1_SCALARS]
                + [(k, ctypes.c_void_p) for k in P1_GEOM_FIELDS + P1_STATE_FIELDS]
                + [("V_min_local", ctypes.c_void_p), ("V_max_local", ctypes.c_void_p)])


def p1_dims(N, NR, g):
    nz, nzr = N + 2 * g, NR + 2 * g
    nf, nfr = nz + 1, nzr + 1
    return dict(zone=nz * nz * nzr, xf=nf * nz * nzr, zf=nz * nz * nfr, fc=nf * nf * nfr)


def p1_struct(prob, arrays, addr=None):
    """A P1Patch over the arrays of a problem; addr(array) gives each array's
    address (host numpy by default)."""
    addr = addr or (lambda a: a.ctypes.data)
    s = P1Patch()
    s.ntp, s.nr, s.nghost = prob["N"], prob["NR"], prob["g"]
    for q in range(4):
        s.itri[q] = 1
    for k in P1_SCALARS:
        setattr(s, k, prob[k])
    for k in P1_GEOM_FIELDS + P1_STATE_FIELDS:
        setattr(s, k, addr(arrays[k]))
    s.V_min_local = arrays["V_min_local"].ctypes.data
    s.V_max_local = arrays["V_max_local"].ctypes.data
    return s


def p1_problem(N, NR, g=4, sector=0, seed=0, shock=True, rmin=1.0, rmax=1.6, gamma=5.0 / 3.0, junk=True,
               noise=0.0):
    """One sector of an (N, NR, g) shell as a patch: the quadrature geometry of
    the oracle (pinned to the reference's own routines), the stencil table,
    and a state with smooth flow, a radial-velocity jump and a pressure jump
    (so the flattener acts); noise > 0 multiplies density, pressure and
    entropy by log-normal zone noise (so the positivity limiter acts).  Output arrays start
    filled with seeded junk so that untouched entries are compared too."""
    import cubesphere_amd as cs  # Config layout only
    cfg = cs.make_config(N, NR, g, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, rmin=rmin, rmax=rmax,
                         gamma=gamma)
    geo = quad_geometry(cfg)
    gd = geometry_dims(cfg)
    sl = lambda a, n, m: np.ascontiguousarray(a[sector * n * m:(sector + 1) * n * m])
    d = p1_dims(N, NR, g)
    st = np.zeros((N + 2 * g) ** 2 * 24, np.int32)
    L = olib()
    L.or_stencil_indexing.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.or_stencil_indexing(N, g, P(st))
    A = dict(vert=sl(geo["vert"], gd["vert"], 3), mask=sl(geo["mask"], gd["zone"], 1), stencil=st,
             scale=sl(geo["scale"], gd["zone"], 1), v_centroid=sl(geo["v_cen"], gd["zone"], 3))
    for f, n in (("xf", "xf"), ("yf", "yf"), ("zf", "zf")):
        nn = gd[n]
        A[f + "_normal"] = sl(geo[f]["normal"], nn, 15 if f == "zf" else 3)
        A[f + "_area"] = sl(geo[f]["area"], nn, 1)
        A[f + "_centroid"] = sl(geo[f]["centroid"], nn, 3)
        b = f[0] + "b"
        A[b + "_pts"] = sl(geo[b]["pts"], nn, 15)
        A[b + "_wts"] = sl(geo[b]["wts"], nn, 5)
        A[b + "_normal"] = sl(geo[b]["normal"], nn, 15 if f == "zf" else 3)
    # arrays MAKE_FACIAL_FLUXES reads besides: area quadrature points, z-face
    # normals (per point), volumes
    for f in ("xf", "yf", "zf"):
        A[f + "_pts"] = sl(geo[f]["pts"], gd[f], 15)
    A["v_volume"] = sl(geo["v_vol"], gd["zone"], 1)
    rng = np.random.default_rng(seed)
    nzone, nfc = d["zone"], d["fc"]
    cen = A["v_centroid"].reshape(3, nzone)
    r = np.sqrt((cen ** 2).sum(0))
    x, y, z = cen / np.maximum(r, 1e-30)
    rho = 1.0 + 0.3 * np.sin(3.0 * x + 1.0) * np.cos(2.0 * y) + 0.2 * (r - rmin)
    prs = 1.0 + 0.2 * np.cos(2.5 * z - 0.5 * x)
    vel = np.stack([0.3 * y - 0.2 * z, -0.3 * x + 0.1 * z, 0.2 * x * y])
    if shock:
        # a front across the middle of the patch: s = r_hat . t, t tangent to the
        # patch centre; converging flow within a band around it
        act = A["mask"] == 1
        c0 = np.array([x[act].mean(), y[act].mean(), z[act].mean()])
        c0 /= np.linalg.norm(c0)
        t = np.cross(c0, [0.3, 0.5, 0.8])
        t /= np.linalg.norm(t)
        sfr = x * t[0] + y * t[1] + z * t[2] - 0.1
        front = sfr > 0.0
        rho = np.where(front, 3.5 * rho, rho)
        prs = np.where(front, 12.0 * prs, prs)
        vel = vel - 2.5 * np.where(front, 1.0, -1.0)[None] * t[:, None] * (np.abs(sfr) < 0.12)[None]
    if noise > 0.0:
        rho = rho * np.exp(noise * rng.normal(size=nzone))
        prs = prs * np.exp(noise * rng.normal(size=nzone))
    Bc = np.stack([0.5 + 0.2 * y, -0.3 + 0.1 * z * x, 0.2 * np.cos(2 * x)])
    Ub = np.zeros((8, nzone))
    Ub[0] = rho
    Ub[1:4] = rho * vel
    Ub[4] = prs / (gamma - 1.0) + 0.5 * rho * (vel ** 2).sum(0) + (Bc ** 2).sum(0) / (8.0 * np.pi)
    Ub[5:8] = Bc
    U = rng.normal(size=(4, 8, nzone))
    U[0] = Ub
    ent = rng.normal(size=(4, nzone))
    ent[0] = prs / rho ** (gamma - 1.0) * (1.0 + 0.01 * rng.uniform(-1, 1, nzone))
    if noise > 0.0:
        ent[0] = ent[0] * np.exp(noise * rng.normal(size=nzone))
    A["U"] = U.ravel().copy()
    A["ent"] = ent.ravel().copy()
    for q, nm in enumerate(("Bxf", "Byf", "Bzf")):
        A[nm] = (0.4 * np.sin(np.arange(nfc) * (0.37 + 0.11 * q)) + 0.1 * q).astype(np.float64)
    sizes = dict(V=32 * nzone, B1_zc=4 * nzone, B2_zc=4 * nzone, B3_zc=4 * nzone, Ex_zc=4 * nzone,
                 Ey_zc=4 * nzone, Ez_zc=4 * nzone, gamma_eos=nzone, gamma_sound=nzone, mean_mol_wt=nzone,
                 divvel=nzone, msonic=nzone, flatten=nzone, work=nzone, Bxf_q=5 * nfc, Byf_q=5 * nfc,
                 Bzf_q=5 * nfc)
    for k, n in sizes.items():
        A[k] = rng.normal(size=n) if junk else np.zeros(n)
    A["V_min_local"] = np.zeros(9)
    A["V_max_local"] = np.zeros(9)
    prob = dict(N=N, NR=NR, g=g, gamma=gamma, molewt=1.0, rhofloor=1e-12, prsfloor=1e-9, smallnum=1e-12,
                flatten_coef=0.4, prs_flatten_coef=1.5)
    return prob, A


def p1_copy(A):
    return {k: v.copy() for k, v in A.items()}


def p1_oracle(prob, A):
    """or_limit_all_slopes_p1 in place on the arrays A."""
    L = olib()
    L.or_limit_all_slopes_p1.argtypes = [ctypes.c_void_p]
    s = p1_struct(prob, A)
    rc = L.or_limit_all_slopes_p1(ctypes.byref(s))
    if rc != 0:
        raise RuntimeError("oracle rejected the patch")
    return A


def p1_reference(prob, A, lib=None):
    """The reference's LIMIT_ALL_SLOPES_P1 (flang build, or the library given)
    in place on A."""
    R = lib if lib is not None else rlib()
    if R is None:
        raise RuntimeError("reference build not present")
    R.ref_limit_all_slopes_p1.argtypes = [ctypes.c_void_p]
    s = p1_struct(prob, A)
    R.ref_limit_all_slopes_p1(ctypes.byref(s))
    return A


# ---------------------------------------------------------------------------
# MAKE_FACIAL_FLUXES of one patch (or_make_facial_fluxes_p1 /
# cubesphere_amd.make_facial_fluxes_p1 share this structure layout)

P1F_SCALARS = ("gamma", "rhofloor", "prsfloor", "smallnum")
P1F_IN = ("vert", "mask", "stencil", "xf_pts", "xf_normal", "xf_area", "xf_centroid", "yf_pts", "yf_normal",
          "yf_area", "yf_centroid", "zf_pts", "zf_normal", "zf_area", "zf_centroid", "scale", "v_volume",
          "v_centroid", "V", "ent", "flatten")
P1F_OUT = ("S_L", "S_R", "S_D", "S_U", "S_B", "S_T", "Fluct_x_Pl", "Fluct_y_Pl", "Fluct_z_Pl", "Fluct_x_Mi",
           "Fluct_y_Mi", "Fluct_z_Mi", "dU_dt", "EntFlux_x", "EntFlux_y", "EntFlux_z", "Efx", "Efy", "Efz")


class P1Flux(ctypes.Structure):
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("pad", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in P1F_SCALARS]
                + [(k, ctypes.c_void_p) for k in P1F_IN + P1F_OUT])


def p1f_struct(prob, arrays, addr=None):
    addr = addr or (lambda a: a.ctypes.data)
    s = P1Flux()
    s.ntp, s.nr, s.nghost = prob["N"], prob["NR"], prob["g"]
    for k in P1F_SCALARS:
        setattr(s, k, prob[k])
    for k in P1F_IN + P1F_OUT:
        setattr(s, k, addr(arrays[k]))
    return s


def p1_flux_problem(N, NR, g=4, sector=0, seed=0, shock=True, noise=0.0, junk=True):
    """A patch reconstructed by the oracle's LIMIT_ALL_SLOPES_P1, with the
    output arrays of MAKE_FACIAL_FLUXES filled with seeded junk."""
    prob, A = p1_problem(N, NR, g, sector=sector, seed=seed, shock=shock, noise=noise)
    p1_oracle(prob, A)
    d = p1_dims(N, NR, g)
    rng = np.random.default_rng(seed + 100)
    sizes = {k: d["fc"] for k in ("S_L", "S_R", "S_D", "S_U", "S_B", "S_T", "EntFlux_x", "EntFlux_y",
                                   "EntFlux_z", "Efx", "Efy", "Efz")}
    for k in ("Fluct_x_Pl", "Fluct_y_Pl", "Fluct_z_Pl", "Fluct_x_Mi", "Fluct_y_Mi", "Fluct_z_Mi"):
        sizes[k] = 8 * d["fc"]
    sizes["dU_dt"] = 8 * d["zone"]
    for k, n in sizes.items():
        A[k] = rng.normal(size=n) if junk else np.zeros(n)
    return prob, A


def p1f_oracle(prob, A):
    L = olib()
    L.or_make_facial_fluxes_p1.argtypes = [ctypes.c_void_p]
    s = p1f_struct(prob, A)
    if L.or_make_facial_fluxes_p1(ctypes.byref(s)) != 0:
        raise RuntimeError("oracle rejected the patch")
    return A



# ---------------------------------------------------------------------------
# UPDATE_IMEX_RK_STEP_P1 of one patch (or_update_imex_rk_step_p1 /
# cubesphere_amd.update_imex_rk_step_p1 share this structure layout)

P1U_SCALARS = ("dtmove", "gamma", "rhofloor", "prsfloor")
P1U_ARRAYS = ("vert", "xf_area", "yf_area", "zf_area", "v_volume", "U", "ent", "Bxf", "Byf", "Bzf", "dU_dt",
              "EntFlux_x", "EntFlux_y", "EntFlux_z", "Efx", "Efy", "Efz", "flag_entprs")


class P1Update(ctypes.Structure):
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("pad", ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in P1U_SCALARS]
                + [(k, ctypes.c_void_p) for k in P1U_ARRAYS])


def p1_update_problem(N, NR, g=4, sector=0, seed=0, dt=1e-3, noise=0.0):
    """A patch after the oracle's LIMIT_ALL_SLOPES_P1 and MAKE_FACIAL_FLUXES,
    with seeded edge electric fields and a time step."""
    prob, A = p1_flux_problem(N, NR, g, sector=sector, seed=seed, noise=noise)
    p1f_oracle(prob, A)
    rng = np.random.default_rng(seed + 200)
    d = p1_dims(N, NR, g)
    for k in ("Efx", "Efy", "Efz"):
        A[k] = rng.normal(0.0, 0.05, d["fc"])
    A["flag_entprs"] = rng.normal(size=d["zone"])
    prob = dict(prob, dtmove=dt)
    return prob, A


def p1u_struct(prob, arrays):
    s = P1Update()
    s.ntp, s.nr, s.nghost = prob["N"], prob["NR"], prob["g"]
    for k in P1U_SCALARS:
        setattr(s, k, prob[k])
    for k in P1U_ARRAYS:
        setattr(s, k, arrays[k].ctypes.data)
    return s


def p1u_oracle(prob, A):
    L = olib()
    L.or_update_imex_rk_step_p1.argtypes = [ctypes.c_void_p]
    s = p1u_struct(prob, A)
    if L.or_update_imex_rk_step_p1(ctypes.byref(s)) != 0:
        raise RuntimeError("oracle rejected the patch")
    return A


# ---------------------------------------------------------------------------
# CLEANUP_AFTER_UPDATE and EVALUATE_TIMESTEP of one patch, on the P1Patch
# structure (cubesphere_amd.cleanup_after_update_p1 / evaluate_timestep_p1)

def p1_cleanup_problem(N, NR, g=4, sector=0, seed=0, dt=2e-3, noise=0.0, floors=True):
    """A patch after the oracle's reconstruction, facial fluxes and update;
    with floors, a few owned zones get a negative internal energy or density
    so that CONS_TO_PRIM takes its entropy and floor branches."""
    prob, A = p1_update_problem(N, NR, g, sector=sector, seed=seed, dt=dt, noise=noise)
    p1u_oracle(prob, A)
    d = p1_dims(N, NR, g)
    nz = d["zone"]
    if floors:
        rng = np.random.default_rng(seed + 300)
        U = A["U"].reshape(4, 8, nz)
        own = np.flatnonzero(A["mask"] == 1)
        pick = rng.choice(own, size=max(2, own.size // 50), replace=False)
        half = pick.size // 2
        U[0, 4, pick[:half]] = 0.5 * (U[0, 1:4, pick[:half]] ** 2).sum(-1) / U[0, 0, pick[:half]] * 0.9
        # a negative density, with a momentum that keeps the velocity O(1) once
        # the density is floored (else the pressure is the difference of two
        # 1e11 energies: rounding noise)
        U[0, 0, pick[half:half + 2]] = -1e-3
        U[0, 1:4, pick[half:half + 2]] *= prob["rhofloor"]
    return prob, A


def p1c_oracle(prob, A):
    L = olib()
    L.or_cleanup_after_update_p1.argtypes = [ctypes.c_void_p]
    s = p1_struct(prob, A)
    if L.or_cleanup_after_update_p1(ctypes.byref(s)) != 0:
        raise RuntimeError("oracle rejected the patch")
    return A


def p1c_reference(prob, A):
    """The reference's CLEANUP_AFTER_UPDATE (flang build) in place on A."""
    R = rlib()
    if R is None:
        raise RuntimeError("reference build not present")
    R.ref_cleanup_after_update_p1.argtypes = [ctypes.c_void_p]
    s = p1_struct(prob, A)
    R.ref_cleanup_after_update_p1(ctypes.byref(s))
    return A


def p1_timestep_oracle(prob, A, cfl, dtcur):
    L = olib()
    L.or_evaluate_timestep_p1.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    s = p1_struct(prob, A)
    out = ctypes.c_double(0.0)
    if L.or_evaluate_timestep_p1(ctypes.byref(s), cfl, dtcur, ctypes.byref(out)) != 0:
        raise RuntimeError("oracle rejected the patch")
    return out.value


def p1_timestep_reference(prob, A, cfl, dtcur):
    """The reference's EVALUATE_TIMESTEP (flang build) in place on A."""
    R = rlib()
    if R is None:
        raise RuntimeError("reference build not present")
    R.ref_evaluate_timestep_p1.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    s = p1_struct(prob, A)
    out = ctypes.c_double(0.0)
    R.ref_evaluate_timestep_p1(ctypes.byref(s), cfl, dtcur, ctypes.byref(out))
    return out.value


# ---------------------------------------------------------------------------
# MAKE_EDGE_EFIELDS_P1 of one patch (or_make_edge_efields_p1 /
# cubesphere_amd.make_edge_efields_p1 share this structure layout)

P1E_IN = ("vert", "mask", "Ex_zc", "Ey_zc", "Ez_zc", "Bxf_q", "Byf_q", "Bzf_q", "S_L", "S_R", "S_D", "S_U", "S_B",
          "S_T")
P1E_OUT = ("Efx", "Efy", "Efz")


class P1Edge(ctypes.Structure):
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("pad", ctypes.c_int32), ("itri", ctypes.c_int32 * 4)]
                + [(k, ctypes.c_void_p) for k in P1E_IN + P1E_OUT])


def p1e_struct(prob, A, itri=(1, 1, 1, 1)):
    s = P1Edge()
    s.ntp, s.nr, s.nghost = prob["N"], prob["NR"], prob["g"]
    for q in range(4):
        s.itri[q] = int(itri[q])
    for k in P1E_IN + P1E_OUT:
        setattr(s, k, A[k].ctypes.data)
    return s


def p1_edge_problem(N, NR, g=4, sector=0, seed=0, shock=True, noise=0.0):
    """A patch after the oracle's LIMIT_ALL_SLOPES_P1 (zone E, face-point B)
    and MAKE_FACIAL_FLUXES (the face wave speeds); the edge arrays hold junk
    that the routine must clear."""
    prob, A = p1_flux_problem(N, NR, g, sector=sector, seed=seed, shock=shock, noise=noise)
    p1f_oracle(prob, A)
    rng = np.random.default_rng(seed + 400)
    for k in P1E_OUT:
        A[k] = rng.normal(size=A[k].size)
    return prob, A


def p1e_oracle(prob, A, itri=(1, 1, 1, 1)):
    L = olib()
    L.or_make_edge_efields_p1.argtypes = [ctypes.c_void_p]
    s = p1e_struct(prob, A, itri)
    if L.or_make_edge_efields_p1(ctypes.byref(s)) != 0:
        raise RuntimeError("oracle rejected the patch")
    return A


def p1e_reference(prob, A, itri=(1, 1, 1, 1)):
    """The reference's MAKE_EDGE_EFIELDS_P1 (flang build) in place on A."""
    R = rlib()
    if R is None:
        raise RuntimeError("reference build not present")
    R.ref_make_edge_efields_p1.argtypes = [ctypes.c_void_p]
    s = p1e_struct(prob, A, itri)
    R.ref_make_edge_efields_p1(ctypes.byref(s))
    return A


# ---------------------------------------------------------------------------
# BUILD_NON_STIFF_SOURCES of one patch (or_build_non_stiff_sources_p1 /
# cubesphere_amd.build_non_stiff_sources_p1 share this structure layout)

class P1Source(ctypes.Structure):
    _fields_ = [("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("newtonsc", ctypes.c_double), ("mstar", ctypes.c_double),
                ("smallnum", ctypes.c_double), ("omega", ctypes.c_double * 3), ("mask", ctypes.c_void_p),
                ("v_centroid", ctypes.c_void_p), ("V", ctypes.c_void_p), ("Source", ctypes.c_void_p),
                ("dU_dt", ctypes.c_void_p)]


P1S_ARRAYS = ("mask", "v_centroid", "V", "Source", "dU_dt")


def p1_source_problem(N, NR, g=4, sector=0, seed=0, gm=(1.0, 0.8), omega=(0.1, -0.2, 0.3)):
    """A reconstructed patch (the oracle's LIMIT_ALL_SLOPES_P1) with seeded
    dU_dt and junk in Source; gm = (newtonsc, Mstar)."""
    prob, A = p1_problem(N, NR, g, sector=sector, seed=seed, shock=True, noise=0.3)
    p1_oracle(prob, A)
    d = p1_dims(N, NR, g)
    rng = np.random.default_rng(seed + 500)
    A["Source"] = rng.normal(size=8 * d["zone"])
    A["dU_dt"] = rng.normal(size=8 * d["zone"])
    prob = dict(prob, newtonsc=gm[0], mstar=gm[1], omega=tuple(omega))
    return prob, A


def p1s_struct(prob, A):
    s = P1Source()
    s.ntp, s.nr, s.nghost = prob["N"], prob["NR"], prob["g"]
    s.newtonsc, s.mstar, s.smallnum = prob["newtonsc"], prob["mstar"], prob["smallnum"]
    for q in range(3):
        s.omega[q] = prob["omega"][q]
    for k in P1S_ARRAYS:
        setattr(s, k, A[k].ctypes.data)
    return s


def p1s_oracle(prob, A):
    L = olib()
    L.or_build_non_stiff_sources_p1.argtypes = [ctypes.c_void_p]
    if L.or_build_non_stiff_sources_p1(ctypes.byref(p1s_struct(prob, A))) != 0:
        raise RuntimeError("oracle rejected the patch")
    return A


# ---------------------------------------------------------------------------
# PAD_BOUNDARY_R of one patch (or_pad_boundary_r_p1 /
# cubesphere_amd.pad_boundary_r_p1 share this structure layout)

class P1Pad(ctypes.Structure):
    _fields_ = [("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("mask", ctypes.c_void_p), ("bcarr_r", ctypes.c_void_p),
                ("xf_area", ctypes.c_void_p), ("yf_area", ctypes.c_void_p), ("zf_area", ctypes.c_void_p),
                ("v_volume", ctypes.c_void_p), ("U", ctypes.c_void_p), ("Bxf", ctypes.c_void_p),
                ("Byf", ctypes.c_void_p), ("Bzf", ctypes.c_void_p)]


P1P_ARRAYS = ("mask", "bcarr_r", "xf_area", "yf_area", "zf_area", "v_volume", "U", "Bxf", "Byf", "Bzf")


def p1_bcarr_r(N, g, inner=2, outer=2):
    """bcarr_r (1-g:N+g, 1-g:N+g, 2) with one code per side; the driver's
    single radial patch pads both sides with 2 (DRV:2238-2247)."""
    nz = N + 2 * g
    b = np.empty(2 * nz * nz, dtype=np.int32)
    b[:nz * nz], b[nz * nz:] = inner, outer
    return b


def p1_pad_problem(N, NR, g=4, sector=0, seed=0, inner=2, outer=2):
    """A patch with seeded state everywhere (so every pad write is visible)
    and bcarr_r set to (inner, outer) over the whole column range."""
    prob, A = p1_problem(N, NR, g, sector=sector, seed=seed, shock=True, noise=0.3)
    rng = np.random.default_rng(seed + 700)
    for k in ("U", "Bxf", "Byf", "Bzf"):
        A[k] = A[k] + 0.01 * rng.normal(size=A[k].size)
    A["bcarr_r"] = p1_bcarr_r(N, g, inner, outer)
    return prob, A


def p1p_struct(prob, A):
    s = P1Pad()
    s.ntp, s.nr, s.nghost = prob["N"], prob["NR"], prob["g"]
    for k in P1P_ARRAYS:
        setattr(s, k, A[k].ctypes.data)
    return s


def p1p_oracle(prob, A):
    L = olib()
    L.or_pad_boundary_r_p1.argtypes = [ctypes.c_void_p]
    if L.or_pad_boundary_r_p1(ctypes.byref(p1p_struct(prob, A))) != 0:
        raise RuntimeError("oracle rejected the patch")
    return A


def p1p_reference(prob, A):
    """The reference's PAD_BOUNDARY_R (flang build) in place on A."""
    R = rlib()
    if R is None:
        raise RuntimeError("reference build not present")
    R.ref_pad_boundary_r_p1.argtypes = [ctypes.c_void_p]
    R.ref_pad_boundary_r_p1(ctypes.byref(p1p_struct(prob, A)))
    return A


# ---------------------------------------------------------------------------
# the driver's time step over the six P1 patches of one shell, on the oracle

class _Stacked:
    def __init__(self, arrs):
        self.a = arrs

    def arrays(self):
        return self.a


def p1_shell_problem(N, NR, g=4, seed=0, noise=0.0, rmin=1.0, rmax=1.6):
    """Six P1 patches (one per sector) of one shell with a smooth state that
    is one function of position (so the ghost zones agree with the
    neighbouring sectors' zones), plus every work array of the stage."""
    import cubesphere_amd as cs  # Config layout only
    probs, As = [], []
    d = p1_dims(N, NR, g)
    for s in range(6):
        prob, A = p1_problem(N, NR, g, sector=s, seed=seed, shock=False, rmin=rmin, rmax=rmax, junk=False,
                             noise=noise)
        for k in P1F_OUT:
            A[k] = np.zeros((8 if k.startswith("Fluct") else 1) * d["fc"])
        A["dU_dt"] = np.zeros(8 * d["zone"])
        A["flag_entprs"] = np.zeros(d["zone"])
        A["Source"] = np.zeros(8 * d["zone"])
        A["gamma_sound"][:] = prob["gamma"]  # as CLEANUP_AFTER_UPDATE leaves it
        probs.append(prob)
        As.append(A)
    cfg = cs.make_config(N, NR, g, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, rmin=rmin, rmax=rmax,
                         radial_bc=cs.RBC_PAD)
    return cfg, probs, As


def p1_shell_exchange_oracle(grid, As):
    nz = As[0]["ent"].size // 4
    keys = (("U", 8 * nz), ("ent", nz), ("Bxf", As[0]["Bxf"].size), ("Byf", As[0]["Byf"].size),
            ("Bzf", As[0]["Bzf"].size))
    st = _Stacked([np.concatenate([A[k][:n] for A in As]) for k, n in keys])
    grid.halo(st)
    for q, (k, n) in enumerate(keys):
        for s, A in enumerate(As):
            A[k][:n] = st.a[q][s * n:(s + 1) * n]


_FAM = ("x", "y", "z")


def p1_shell_sync_oracle(As, faces, edges):
    """The driver's defect correction in its before-the-update form: the
    fluctuations and entropy flux of each shared face become the mean of the
    two patches' (334fetch_buffers_fill_in_flux_x.f:100-116, 341...f), the
    patch's boundary zone takes -(mean - own) A / V in dU_dt (485update_defect
    _imex_rk.f), each shared edge's field the mean of its two or three
    patches' (334...f:128-160, 340...f).  faces, edges: the pairings of
    cubesphere_amd.p1_shell_pairings."""
    nz = As[0]["ent"].size // 4
    nfc = As[0]["Bxf"].size
    for q in faces:
        fa, fb = _FAM[q["fa"]], _FAM[q["fb"]]
        A, B = As[q["pa"]], As[q["pb"]]
        ia, ib = int(q["ia"]), int(q["ib"])
        sl = lambda i: slice(i, i + 8 * nfc, nfc)
        Pa, Ma, Ea = A["Fluct_%s_Pl" % fa][sl(ia)].copy(), A["Fluct_%s_Mi" % fa][sl(ia)].copy(), A["EntFlux_" + fa][ia]
        Pb, Mb, Eb = B["Fluct_%s_Pl" % fb][sl(ib)].copy(), B["Fluct_%s_Mi" % fb][sl(ib)].copy(), B["EntFlux_" + fb][ib]
        flip = bool(q["flip"])
        mP = 0.5 * (Pa + (Mb if flip else Pb))
        mM = 0.5 * (Ma + (Pb if flip else Mb))
        mE = 0.5 * (Ea + (-Eb if flip else Eb))
        for X, fam, i, z, ar, side, own_P, own_M, nP, nM, nE in (
                (A, fa, ia, int(q["za"]), int(q["aa"]), int(q["sa"]), Pa, Ma, mP, mM, mE),
                (B, fb, ib, int(q["zb"]), int(q["ab"]), int(q["sb"]), Pb, Mb, mM if flip else mP, mP if flip else mM,
                 -mE if flip else mE)):
            w = X[fam + "f_area"][ar] / X["v_volume"][z]
            own, mean = (own_P, nP) if side else (own_M, nM)
            X["dU_dt"][z::nz][:8] += -(mean - own) * w
            X["Fluct_%s_Pl" % fam][sl(i)] = nP
            X["Fluct_%s_Mi" % fam][sl(i)] = nM
            X["EntFlux_" + fam][i] = nE
    for q in edges:
        n = int(q["n"])
        tot = 0.0
        for m in range(n):
            tot += q["sign"][m] * As[q["p"][m]]["Ef" + _FAM[q["f"][m]]][q["idx"][m]]
        mean = tot * (1.0 / 3.0) if n == 3 else 0.5 * tot
        for m in range(n):
            As[q["p"][m]]["Ef" + _FAM[q["f"][m]]][q["idx"][m]] = q["sign"][m] * mean


def p1_shell_step_oracle(grid, probs, As, dt, sources=(1.0, 0.0, (0.0, 0.0, 0.0)), pairings=None, pad=True):
    """One Heun RK2 step of the driver (DRV:1746-3461) over the six patches,
    every routine the oracle's; with pairings (faces, edges) the defect
    correction in its before-the-update form; with pad the outflow
    PAD_BOUNDARY_R at both radial ends after each exchange (DRV:2238-2275)."""
    nz = As[0]["ent"].size // 4
    keys = (("U", 8 * nz), ("ent", nz), ("Bxf", As[0]["Bxf"].size), ("Byf", As[0]["Byf"].size),
            ("Bzf", As[0]["Bzf"].size))
    saved = [None] * 6
    for irk in (1, 2):
        p1_shell_exchange_oracle(grid, As)
        if pad:
            for prob, A in zip(probs, As):
                p1p_oracle(prob, dict(A, bcarr_r=p1_bcarr_r(prob["N"], prob["g"])))
        for s, (prob, A) in enumerate(zip(probs, As)):
            p1_oracle(prob, A)
            p1f_oracle(prob, A)
            p1e_oracle(prob, A)
            p1s_oracle(dict(prob, newtonsc=sources[0], mstar=sources[1], omega=tuple(sources[2])), A)
        if pairings is not None:
            p1_shell_sync_oracle(As, *pairings)
        for s, (prob, A) in enumerate(zip(probs, As)):
            if irk == 1:
                saved[s] = {k: A[k][:n].copy() for k, n in keys}
                move = dt
            else:
                for k, n in keys:
                    A[k][:n] = (A[k][:n] + saved[s][k]) * 0.5
                move = 0.5 * dt
            p1u_oracle(dict(prob, dtmove=move), A)
            p1c_oracle(prob, A)
    return As


# ---------------------------------------------------------------------------
# The P1 stage of a shell (the product's CS_RECON_P1ADP): the reference's
# routines (the oracle's LIMIT_ALL_SLOPES_P1, MAKE_EDGE_EFIELDS_P1,
# CLEANUP_AFTER_UPDATE, EVALUATE_TIMESTEP), one resolved flux per face
# (or_p1s_face_fluxes), the panel-edge averages before the update
# (or_sync_shared), the stage update (or_p1s_update_zones and the arc-length
# constrained transport of or_update_imex_rk_step_p1), the geometric exchange.

def p1s_patches(cfg):
    """Six per-sector P1 patches holding the quadrature geometry (pinned to
    the reference's routines) and zeroed work arrays; the state lives in the
    stage layout and is copied in and out per stage."""
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    geo = quad_geometry(cfg)
    gd = geometry_dims(cfg)
    d = p1_dims(N, NR, g)
    st = np.zeros((N + 2 * g) ** 2 * 24, np.int32)
    L = olib()
    L.or_stencil_indexing.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.or_stencil_indexing(N, g, P(st))
    As = []
    for s in range(6):
        sl = lambda a, n, m: np.ascontiguousarray(a[s * n * m:(s + 1) * n * m])
        A = dict(vert=sl(geo["vert"], gd["vert"], 3), mask=sl(geo["mask"], gd["zone"], 1), stencil=st,
                 scale=sl(geo["scale"], gd["zone"], 1), v_centroid=sl(geo["v_cen"], gd["zone"], 3),
                 v_volume=sl(geo["v_vol"], gd["zone"], 1))
        for f in ("xf", "yf", "zf"):
            nn = gd[f]
            A[f + "_normal"] = sl(geo[f]["normal"], nn, 15 if f == "zf" else 3)
            A[f + "_area"] = sl(geo[f]["area"], nn, 1)
            A[f + "_centroid"] = sl(geo[f]["centroid"], nn, 3)
            A[f + "_pts"] = sl(geo[f]["pts"], nn, 15)
            b = f[0] + "b"
            A[b + "_pts"] = sl(geo[b]["pts"], nn, 15)
            A[b + "_wts"] = sl(geo[b]["wts"], nn, 5)
            A[b + "_normal"] = sl(geo[b]["normal"], nn, 15 if f == "zf" else 3)
        nz, nfc = d["zone"], d["fc"]
        sizes = dict(U=32 * nz, ent=4 * nz, Bxf=nfc, Byf=nfc, Bzf=nfc, V=32 * nz, B1_zc=4 * nz, B2_zc=4 * nz,
                     B3_zc=4 * nz, Ex_zc=4 * nz, Ey_zc=4 * nz, Ez_zc=4 * nz, gamma_eos=nz, gamma_sound=nz,
                     mean_mol_wt=nz, divvel=nz, msonic=nz, flatten=nz, work=nz, Bxf_q=5 * nfc, Byf_q=5 * nfc,
                     Bzf_q=5 * nfc, dU_dt=8 * nz, flag_entprs=nz)
        for k, n in sizes.items():
            A[k] = np.zeros(n)
        for k in P1F_OUT:
            if k not in A:
                A[k] = np.zeros((8 if k.startswith("Fluct") else 1) * nfc)
        A["V_min_local"] = np.zeros(9)
        A["V_max_local"] = np.zeros(9)
        As.append(A)
    return As


def _p1s_prob(cfg):
    return dict(N=cfg.ntp, NR=cfg.nr, g=cfg.nghost, gamma=cfg.gamma, molewt=1.0, rhofloor=cfg.rhofloor,
                prsfloor=cfg.prsfloor, smallnum=1e-12, flatten_coef=0.4, prs_flatten_coef=1.5)


def _p1s_radial_boundary(cfg, As, st):
    """the radial ghost layers after the exchange: PAD_BOUNDARY_R with outflow
    at both ends (shells with radial_bc 0) or ADD_BOUNDARY with the problem
    bc_problem (radial_bc 2)"""
    if cfg.nr < 2 or cfg.radial_bc == 1:
        return
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    arrs = st.arrays()
    if cfg.radial_bc == 2:
        new = init_problem_p1(cfg, cfg.bc_problem, layers=3, state=arrs)
        for a, b in zip(arrs, new):
            a[:] = b
        return
    nz = (N + 2 * g) ** 2 * (NR + 2 * g)
    nfc = (N + 2 * g + 1) ** 2 * (NR + 2 * g + 1)
    prob = dict(N=N, NR=NR, g=g)
    for s, A in enumerate(As):
        P = dict(mask=A["mask"], bcarr_r=p1_bcarr_r(N, g), xf_area=A["xf_area"], yf_area=A["yf_area"],
                 zf_area=A["zf_area"], v_volume=A["v_volume"],
                 U=np.ascontiguousarray(arrs[0][s * 8 * nz:(s + 1) * 8 * nz]),
                 Bxf=np.ascontiguousarray(arrs[2][s * nfc:(s + 1) * nfc]),
                 Byf=np.ascontiguousarray(arrs[3][s * nfc:(s + 1) * nfc]),
                 Bzf=np.ascontiguousarray(arrs[4][s * nfc:(s + 1) * nfc]))
        p1p_oracle(prob, P)
        arrs[0][s * 8 * nz:(s + 1) * 8 * nz] = P["U"]
        arrs[2][s * nfc:(s + 1) * nfc] = P["Bxf"]
        arrs[3][s * nfc:(s + 1) * nfc] = P["Byf"]
        arrs[4][s * nfc:(s + 1) * nfc] = P["Bzf"]


def p1s_stage_oracle(grid, cfg, As, st_in, st_u0, st_out, dt, irk):
    """One stage of the P1 stage of a shell on HostState-like objects with
    .arrays() = [U, ent, Bx, By, Bz] in the stage layout.  st_in is modified
    in place (exchange, primitives' floors, the reconstruction's zone B and
    energy), as the device stage modifies its input buffer."""
    L = olib()
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    prob = _p1s_prob(cfg)
    d = p1_dims(N, NR, g)
    nz, nfc = d["zone"], d["fc"]
    order, imex = cfg.rk_order, cfg.imex
    if imex:
        wa, wb = (0.0, 1.0) if irk == 1 else (0.5, 0.5)
        mode = irk
    elif order == 3:
        wa, wb = ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))[irk - 1]
        mode = 0
    else:
        wa, wb = ((0.0, 1.0), (0.5, 0.5))[irk - 1]
        mode = 0
    grid.halo(st_in)
    _p1s_radial_boundary(cfg, As, st_in)
    Ui, Ei, Bi = st_in.arrays()[0], st_in.arrays()[1], st_in.arrays()[2:]
    F = [np.zeros(6 * 9 * nfc) for _ in range(3)]
    S = [np.zeros(6 * 2 * nfc) for _ in range(3)]
    E = [np.zeros(6 * nfc) for _ in range(3)]
    dbl = ctypes.POINTER(ctypes.c_double)
    L.or_p1s_face_fluxes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    for s, A in enumerate(As):
        A["U"][:] = 0.0
        A["U"][:8 * nz] = Ui[s * 8 * nz:(s + 1) * 8 * nz]
        A["ent"][:] = 0.0
        A["ent"][:nz] = Ei[s * nz:(s + 1) * nz]
        for q, k in enumerate(("Bxf", "Byf", "Bzf")):
            A[k][:] = Bi[q][s * nfc:(s + 1) * nfc]
        p1_oracle(prob, A)
        Ui[s * 8 * nz:(s + 1) * 8 * nz] = A["U"][:8 * nz]
        Ei[s * nz:(s + 1) * nz] = A["ent"][:nz]
        for q, k in enumerate(("Bxf", "Byf", "Bzf")):
            Bi[q][s * nfc:(s + 1) * nfc] = A[k]
        Fs = [np.zeros(9 * nfc) for _ in range(3)]
        Ss = [np.zeros(2 * nfc) for _ in range(3)]
        fp = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in Fs])
        sp = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in Ss])
        sf = p1f_struct(prob, A)
        if L.or_p1s_face_fluxes(ctypes.byref(sf), int(cfg.riemann), int(cfg.radial_bc == 1), fp, sp) != 0:
            raise RuntimeError("oracle rejected the patch")
        for q in range(3):
            F[q][s * 9 * nfc:(s + 1) * 9 * nfc] = Fs[q]
            S[q][s * 2 * nfc:(s + 1) * 2 * nfc] = Ss[q]
        A["S_L"][:], A["S_R"][:] = Ss[0][:nfc], Ss[0][nfc:]
        A["S_D"][:], A["S_U"][:] = Ss[1][:nfc], Ss[1][nfc:]
        A["S_B"][:], A["S_T"][:] = Ss[2][:nfc], Ss[2][nfc:]
        p1e_oracle(prob, A)
        for q, k in enumerate(("Efx", "Efy", "Efz")):
            e = A[k].copy()
            if cfg.radial_bc == 1 and q < 2:  # conducting walls: no field on their edges
                e3 = e.reshape(NR + 2 * g + 1, -1)
                e3[g] = 0.0
                e3[g + NR] = 0.0
            E[q][s * nfc:(s + 1) * nfc] = e
    L.or_sync_shared.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.or_sync_shared(grid.g, (ctypes.c_void_p * 3)(*[a.ctypes.data for a in F]),
                     (ctypes.c_void_p * 3)(*[a.ctypes.data for a in E]))
    Uz, Ez, Bz_ = st_u0.arrays()[0], st_u0.arrays()[1], st_u0.arrays()[2:]
    Uo, Eo, Bo = st_out.arrays()[0], st_out.arrays()[1], st_out.arrays()[2:]
    L.or_p1s_update_zones.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_void_p] * 7 + [ctypes.c_double] * 3 +
                                      [ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_void_p] * 2)
    L.or_update_imex_rk_step_p1.argtypes = [ctypes.c_void_p]
    for s, A in enumerate(As):
        # faces: constrained transport of the stage input's faces, then the weights
        W = p1_copy({k: A[k] for k in P1U_ARRAYS})
        W["U"] = np.zeros(32 * nz)
        W["ent"] = np.zeros(4 * nz)
        W["ent"][:nz] = 1.0
        W["U"][:nz] = 1.0
        W["U"][4 * nz:5 * nz] = 1.0
        W["dU_dt"] = np.zeros(8 * nz)
        for k in ("EntFlux_x", "EntFlux_y", "EntFlux_z"):
            W[k] = np.zeros(nfc)
        for q, k in enumerate(("Efx", "Efy", "Efz")):
            W[k] = E[q][s * nfc:(s + 1) * nfc].copy()
        for q, k in enumerate(("Bxf", "Byf", "Bzf")):
            W[k] = Bi[q][s * nfc:(s + 1) * nfc].copy()
        p1u_oracle(dict(prob, dtmove=dt), W)
        # owned faces only: x (0..N, 1..N, 1..NR), y (1..N, 0..N, 1..NR), z (1..N, 1..N, 0..NR)
        nf = N + 2 * g + 1
        for q, k in enumerate(("Bxf", "Byf", "Bzf")):
            i0, j0, k0 = (0 if q == 0 else 1), (0 if q == 1 else 1), (0 if q == 2 else 1)
            new = W[k].reshape(NR + 2 * g + 1, nf, nf)
            b0 = Bz_[q][s * nfc:(s + 1) * nfc].reshape(NR + 2 * g + 1, nf, nf)
            bo = Bo[q][s * nfc:(s + 1) * nfc].reshape(NR + 2 * g + 1, nf, nf)
            sel = (slice(k0 + g, NR + 1 + g), slice(j0 + g, N + 1 + g), slice(i0 + g, N + 1 + g))
            bo[sel] = (wa * b0[sel] + wb * new[sel]) if wa != 0.0 else wb * new[sel]
        # zones
        ptr = lambda a: a.ctypes.data
        fsec = [np.ascontiguousarray(F[q][s * 9 * nfc:(s + 1) * 9 * nfc]) for q in range(3)]
        ui = np.ascontiguousarray(Ui[s * 8 * nz:(s + 1) * 8 * nz])
        ei = np.ascontiguousarray(Ei[s * nz:(s + 1) * nz])
        uz = np.ascontiguousarray(Uz[s * 8 * nz:(s + 1) * 8 * nz])
        ez = np.ascontiguousarray(Ez[s * nz:(s + 1) * nz])
        uo = np.ascontiguousarray(Uo[s * 8 * nz:(s + 1) * 8 * nz])
        eo = np.ascontiguousarray(Eo[s * nz:(s + 1) * nz])
        L.or_p1s_update_zones(N, NR, g, (ctypes.c_void_p * 3)(*[ptr(a) for a in fsec]), ptr(A["v_volume"]),
                              ptr(A["v_centroid"]), ptr(ui), ptr(ei), ptr(uz), ptr(ez), dt, wa, wb, mode,
                              float(cfg.grav), cfg.gamma, cfg.rhofloor, cfg.prsfloor, ptr(uo), ptr(eo))
        Uo[s * 8 * nz:(s + 1) * 8 * nz] = uo
        Eo[s * nz:(s + 1) * nz] = eo
    # CLEANUP_AFTER_UPDATE on the output (the new faces), owned zones kept
    for s, A in enumerate(As):
        C = p1_copy(A)
        C["U"][:] = 0.0
        C["U"][:8 * nz] = Uo[s * 8 * nz:(s + 1) * 8 * nz]
        C["ent"][:] = 0.0
        C["ent"][:nz] = Eo[s * nz:(s + 1) * nz]
        for q, k in enumerate(("Bxf", "Byf", "Bzf")):
            C[k] = Bo[q][s * nfc:(s + 1) * nfc].copy()
        p1c_oracle(prob, C)
        own = np.zeros((NR + 2 * g, N + 2 * g, N + 2 * g), bool)
        own[g:g + NR, g:g + N, g:g + N] = True
        own = own.ravel()
        for q in range(8):
            blk = Uo[s * 8 * nz + q * nz:s * 8 * nz + (q + 1) * nz]
            blk[own] = C["U"][q * nz:(q + 1) * nz][own]
        Eo[s * nz:(s + 1) * nz][own] = C["ent"][:nz][own]


def p1s_cfl_oracle(cfg, As, st, dt_old):
    """EVALUATE_TIMESTEP over the six sectors (U of the owned active zones
    rewritten by CONS_TO_PRIM, as the routine does); the minimum of the six."""
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    prob = _p1s_prob(cfg)
    d = p1_dims(N, NR, g)
    nz = d["zone"]
    U, E = st.arrays()[0], st.arrays()[1]
    best = None
    for s, A in enumerate(As):
        A["U"][:] = 0.0
        A["U"][:8 * nz] = U[s * 8 * nz:(s + 1) * 8 * nz]
        A["ent"][:] = 0.0
        A["ent"][:nz] = E[s * nz:(s + 1) * nz]
        A["gamma_sound"][:] = cfg.gamma
        dtn = p1_timestep_oracle(prob, A, cfg.cfl, dt_old)
        U[s * 8 * nz:(s + 1) * 8 * nz] = A["U"][:8 * nz]
        best = dtn if best is None else min(best, dtn)
    return best


def p1s_step_oracle(grid, cfg, As, st, dt):
    """One step of the P1 stage (stage weights of SSP-RK3, Heun RK2 or
    IMEX-SSP2(2,2,2)) on st in place; returns the next time step."""
    other = st.copy()
    if cfg.imex:
        U, E = st.arrays()[0], st.arrays()[1]
        U1, E1 = other.arrays()[0], other.arrays()[1]
        B, B1 = st.arrays()[2:], other.arrays()[2:]
        N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
        nz = (N + 2 * g) ** 2 * (NR + 2 * g)
        L = olib()
        L.or_imex_pre.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        U1[:] = U
        E1[:] = E
        for q in range(3):
            B1[q][:] = B[q]
        for s, A in enumerate(As):
            mask = A["mask"]
            cen = A["v_centroid"].reshape(3, nz)
            for z in np.nonzero(mask == 1)[0]:
                c = cen[:, z]
                r2 = float(c @ c)
                n = np.ascontiguousarray(c / np.sqrt(r2))
                u = np.ascontiguousarray(U1[s * 8 * nz + z::nz][:8])
                L.or_imex_pre(dt, -cfg.grav / r2, n.ctypes.data, u.ctypes.data)
                U1[s * 8 * nz + z::nz][:8] = u
        p1s_stage_oracle(grid, cfg, As, other, st, other, dt, 1)
        p1s_stage_oracle(grid, cfg, As, other, st, st, dt, 2)
    elif cfg.rk_order == 3:
        p1s_stage_oracle(grid, cfg, As, st, st, other, dt, 1)
        p1s_stage_oracle(grid, cfg, As, other, st, other, dt, 2)
        p1s_stage_oracle(grid, cfg, As, other, st, st, dt, 3)
    else:
        p1s_stage_oracle(grid, cfg, As, st, st, other, dt, 1)
        p1s_stage_oracle(grid, cfg, As, other, st, st, dt, 2)
    return p1s_cfl_oracle(cfg, As, st, dt)


# ---------------------------------------------------------------------------
# PROBLEM_GENERATOR_U_A + INIT_PATCH_CUBE_SPHERE / ADD_BOUNDARY over the six
# P1 patches of a shell (or_init_patch_cube_sphere, or the reference's
# routines built from source: ref_init_patch)

class OrProb(ctypes.Structure):
    _fields_ = [("iproblem", ctypes.c_int32), ("pad", ctypes.c_int32), ("time", ctypes.c_double),
                ("rmin", ctypes.c_double), ("rmax", ctypes.c_double), ("gamma", ctypes.c_double),
                ("rhofloor", ctypes.c_double), ("prsfloor", ctypes.c_double), ("grav", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


P1INIT_ARRAYS = ("vert", "mask", "v_pts", "v_wts", "xf_area", "yf_area", "zf_area", "U", "ent", "Bxf", "Byf",
                 "Bzf")


class OrP1Init(ctypes.Structure):
    _fields_ = ([("ntp", ctypes.c_int32), ("nr", ctypes.c_int32), ("nghost", ctypes.c_int32),
                 ("sector", ctypes.c_int32)] + [(k, ctypes.c_void_p) for k in P1INIT_ARRAYS])


def or_prob(cfg, problem, seed=0, time=0.0):
    return OrProb(int(problem), 0, float(time), float(cfg.rmin), float(cfg.rmax), float(cfg.gamma),
                  float(cfg.rhofloor), float(cfg.prsfloor), float(cfg.grav), int(seed))


def problem_point(cfg, problem, x, seed=0, time=0.0):
    """PROBLEM_GENERATOR_U_A at one point: (U_pt(8), A_pt(3))"""
    L = olib()
    L.or_problem_generator_u_a.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 2
    U, A = np.zeros(8), np.zeros(3)
    pb = or_prob(cfg, problem, seed, time)
    L.or_problem_generator_u_a(ctypes.byref(pb), float(x[0]), float(x[1]), float(x[2]), P(U), P(A))
    return U, A


def init_problem_p1(cfg, problem, seed=0, time=0.0, layers=0, state=None, geo=None, use_reference=False):
    """INIT_PATCH_CUBE_SPHERE (layers 0) or ADD_BOUNDARY (layers 1 inner,
    2 outer, 3 both; on a copy of `state`) over the six sectors on the
    quadrature geometry; returns (U, ent, Bx, By, Bz) in the stage layout
    (sector-major)."""
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    geo = quad_geometry(cfg) if geo is None else geo
    gd = geometry_dims(cfg)
    nz = (N + 2 * g) ** 2 * (NR + 2 * g)
    nfc = (N + 2 * g + 1) ** 2 * (NR + 2 * g + 1)
    if state is None:
        U, ent, Bx, By, Bz = np.zeros(6 * 8 * nz), np.zeros(6 * nz), np.zeros(6 * nfc), np.zeros(6 * nfc), \
            np.zeros(6 * nfc)
    else:
        U, ent, Bx, By, Bz = (np.array(a, dtype=np.float64, copy=True) for a in state)
    pb = or_prob(cfg, problem, seed, time)
    if use_reference:
        R = rlib()
        if R is None:
            raise RuntimeError("reference build not present")
        fn = R.ref_init_patch
    else:
        fn = olib().or_init_patch_cube_sphere
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    for s in range(6):
        sl = lambda a, n, m: np.ascontiguousarray(a[s * n * m:(s + 1) * n * m])
        A = dict(vert=sl(geo["vert"], gd["vert"], 3), mask=sl(geo["mask"], gd["zone"], 1),
                 v_pts=sl(geo["v_pts"], gd["zone"], 24), v_wts=sl(geo["v_wts"], gd["zone"], 8),
                 xf_area=sl(geo["xf"]["area"], gd["xf"], 1), yf_area=sl(geo["yf"]["area"], gd["yf"], 1),
                 zf_area=sl(geo["zf"]["area"], gd["zf"], 1),
                 U=np.ascontiguousarray(U[s * 8 * nz:(s + 1) * 8 * nz]), ent=np.ascontiguousarray(ent[s * nz:(s + 1) * nz]),
                 Bxf=np.ascontiguousarray(Bx[s * nfc:(s + 1) * nfc]), Byf=np.ascontiguousarray(By[s * nfc:(s + 1) * nfc]),
                 Bzf=np.ascontiguousarray(Bz[s * nfc:(s + 1) * nfc]))
        st = OrP1Init(N, NR, g, s, *[A[k].ctypes.data for k in P1INIT_ARRAYS])
        rc = fn(ctypes.byref(st), ctypes.byref(pb), int(layers))
        if rc not in (None, 0):
            raise RuntimeError("init rejected the patch")
        U[s * 8 * nz:(s + 1) * 8 * nz] = A["U"]
        ent[s * nz:(s + 1) * nz] = A["ent"]
        Bx[s * nfc:(s + 1) * nfc] = A["Bxf"]
        By[s * nfc:(s + 1) * nfc] = A["Byf"]
        Bz[s * nfc:(s + 1) * nfc] = A["Bzf"]
    return U, ent, Bx, By, Bz


# ---------------------------------------------------------------------------
# Third order (NUMORDER 3): the 27-point volume rule, ten basis moments and
# LIMIT_ALL_SLOPES_P12 on one patch.

def volume_geometry3(cfg, geo, use_reference=False):
    """COORDS_TO_WT_QUADPTS_VOLUME_ISO2 at NUMORDER 3 (27 points) and
    INIT_BASIS_FUNCTION_COEFS with ten modes, every sector, from the vertices,
    mask and scale lengths of quad_geometry(cfg): v_pts (zone, 3, 27), v_wts
    (zone, 27), v_tan (zone, 3, 3, 27), v_vol, v_cen (zone, 3), moments
    (zone, 10), sector slowest."""
    gd = geometry_dims(cfg)
    nzn = gd["zone"]
    a = dict(v_pts=np.zeros(6 * nzn * 81), v_wts=np.zeros(6 * nzn * 27), v_tan=np.zeros(6 * nzn * 243),
             v_vol=np.zeros(6 * nzn), v_cen=np.zeros(6 * nzn * 3), moments=np.zeros(6 * nzn * 10))
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    vp = ctypes.c_void_p
    if use_reference:
        R = rlib3()
        if R is None:
            raise RuntimeError("third-order reference build not present")
        R.ref3_quad_volume(N, NR, g, P(geo["vert"]), P(geo["mask"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_tan"]),
                           P(a["v_vol"]), P(a["v_cen"]))
        R.ref3_basis_moments(N, NR, g, P(geo["mask"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_vol"]), P(a["v_cen"]),
                             P(geo["scale"]), P(a["moments"]))
        return a
    L = olib()
    L.or_quad_volume3.argtypes = [vp] * 7
    L.or_basis_moments10.argtypes = [vp] * 7
    L.or_quad_volume3(ctypes.byref(cfg), P(geo["vert"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_tan"]), P(a["v_vol"]),
                      P(a["v_cen"]))
    L.or_basis_moments10(ctypes.byref(cfg), P(geo["mask"]), P(a["v_pts"]), P(a["v_wts"]), P(a["v_cen"]),
                         P(geo["scale"]), P(a["moments"]))
    return a


P12_EXTRA = ("xf_pts", "xf_wts", "yf_pts", "yf_wts", "zf_pts", "zf_wts", "zf_normal", "v_volume", "moments", "Wzc")


class P12Extra(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in P12_EXTRA]


def p12_problem(N, NR, g=4, sector=0, seed=0, shock=False, rmin=1.0, rmax=1.6, gamma=5.0 / 3.0, noise=0.0):
    """One sector as a patch for the third-order routine: the arrays of
    p1_problem with ten-mode U, V, B*_zc, E*_zc and Wzc (seeded junk above
    the means), the 27-point centroids and volumes, the ten basis moments,
    and the face quadrature arrays the routine declares."""
    import cubesphere_amd as cs  # Config layout only
    prob, A = p1_problem(N, NR, g, sector=sector, seed=seed, shock=shock, rmin=rmin, rmax=rmax, gamma=gamma,
                         noise=noise)
    cfg = cs.make_config(N, NR, g, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, rmin=rmin, rmax=rmax,
                         gamma=gamma)
    geo = quad_geometry(cfg)
    g3 = volume_geometry3(cfg, geo)
    gd = geometry_dims(cfg)
    d = p1_dims(N, NR, g)
    nzone = d["zone"]
    sl = lambda a, n, m: np.ascontiguousarray(a[sector * n * m:(sector + 1) * n * m])
    rng = np.random.default_rng(seed + 1000)
    U = rng.normal(size=(10, 8 * nzone))
    U[0] = A["U"][:8 * nzone]
    A["U"] = U.ravel().copy()
    A["V"] = rng.normal(size=80 * nzone)
    for k in ("B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc", "Wzc"):
        A[k] = rng.normal(size=10 * nzone)
    A["v_centroid"] = sl(g3["v_cen"], gd["zone"], 3)
    A["v_volume"] = sl(g3["v_vol"], gd["zone"], 1)
    A["moments"] = sl(g3["moments"], gd["zone"], 10)
    for f in ("xf", "yf", "zf"):
        A[f + "_wts"] = sl(geo[f]["wts"], gd[f], 5)
    return prob, A


def p12_oracle(prob, A):
    """or_limit_all_slopes_p12 in place on A (the whole routine)."""
    L = olib()
    L.or_limit_all_slopes_p12.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    s = p1_struct(prob, A)
    if L.or_limit_all_slopes_p12(ctypes.byref(s), P(A["moments"])) != 0:
        raise RuntimeError("oracle rejected the patch")
    return A


def p12_reference(prob, A):
    """The reference's whole LIMIT_ALL_SLOPES_P12 (NUMORDER 3 build) in place
    on A."""
    R = rlib3()
    if R is None:
        raise RuntimeError("third-order reference build not present")
    R.ref3_limit_all_slopes_p12.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    s = p1_struct(prob, A)
    e = P12Extra()
    for k in P12_EXTRA:
        setattr(e, k, A[k].ctypes.data)
    R.ref3_limit_all_slopes_p12(ctypes.byref(s), ctypes.byref(e))
    return A
