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
