"""Generate the golden fixtures in this directory (run in the build container,
where /root/reference and oracle/_ref/libcsref.so exist; the fixtures are
committed and the tests only read them).

ref_vectors.npz   input/output vectors of reference routines compiled from the
                  reference's own sources (oracle/ref/Makefile):
                  Mesh_Sector_General + FILL_GHOST2D_GLOBAL (201, 203),
                  GENERATE_GLOBAL_RADIAL_UNIFORM/LOG_MESH (202),
                  CONS_TO_PRIM / PRIM_TO_CONS / PRIM_TO_WAVSPD (662, 661, 660),
                  Two_D_RS_Ez_4State_HLL / _4State_LLF / _3State_LLF (630, 632, 633).
quadpoints.json   the numbers of _T/_200mesh/test_quadpoints.out, the
                  reference's one committed output file.
"""
import ctypes
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_OUT = "/root/reference/_T/_200mesh/test_quadpoints.out"


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libcsref.so"))
    rng = np.random.default_rng(20261202)
    out = {}
    # lattices
    for N, g in [(8, 4), (12, 3), (17, 3)]:
        nf = N + 2 * g + 1
        a = np.zeros(nf * nf * 3 * 6)
        lib.ref_unit_lattice(ctypes.c_int(N), ctypes.c_int(g), P(a))
        out[f"lattice_N{N}_g{g}"] = a
    # radial meshes
    lib.ref_radial_mesh.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_void_p]
    for nr, g, r0, r1, lg in [(8, 3, 1.0, 1.5, 0), (16, 4, 1.0, 2.0, 1)]:
        rb = np.zeros(nr + 2 * g + 1)
        lib.ref_radial_mesh(nr, g, r0, r1, lg, P(rb))
        out[f"radial_nr{nr}_g{g}_log{lg}"] = rb
        out[f"radial_nr{nr}_g{g}_log{lg}_args"] = np.array([nr, g, r0, r1, lg], float)
    # pointwise PDE routines
    gam, rhof, prsf = 5.0 / 3.0, 1e-12, 1e-9
    n = 400
    V = np.empty((n, 8))
    V[:, 0] = 10 ** rng.uniform(-3, 1, n)
    V[:, 1:4] = rng.normal(0, 2, (n, 3))
    V[:, 4] = 10 ** rng.uniform(-4, 1, n)
    V[:, 5:8] = rng.normal(0, 3, (n, 3))
    V[:20, 4] = -np.abs(V[:20, 4])        # floors and the entropy fallback
    V[20:30, 0] = 1e-14
    ent = 10 ** rng.uniform(-2, 0, n)
    lib.ref_prim_to_cons.argtypes = [ctypes.c_double] * 4 + [ctypes.c_void_p] * 2
    lib.ref_cons_to_prim.argtypes = [ctypes.c_double] * 3 + [ctypes.c_void_p] * 4
    lib.ref_wavspd.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    U = np.zeros((n, 8))
    Vin = V.copy()
    Vout = V.copy()
    for i in range(n):
        v = Vout[i].copy()
        u = np.zeros(8)
        lib.ref_prim_to_cons(gam, rhof, prsf, ent[i], P(v), P(u))
        Vout[i] = v
        U[i] = u
    out["p2c_V_in"], out["p2c_ent"], out["p2c_V_out"], out["p2c_U"] = Vin, ent, Vout, U
    Uin = U.copy()
    Uin[:15, 4] = 0.0                      # negative pressure from energy
    Uout = Uin.copy()
    Vc = np.zeros((n, 8))
    eo = ent.copy()
    fl = np.zeros(n, np.int32)
    for i in range(n):
        u = Uout[i].copy()
        v = np.zeros(8)
        e = np.array([eo[i]])
        f = np.zeros(1, np.int32)
        lib.ref_cons_to_prim(gam, rhof, prsf, P(e), P(u), P(v), P(f))
        Uout[i], Vc[i], eo[i], fl[i] = u, v, e[0], f[0]
    out["c2p_U_in"], out["c2p_ent_in"], out["c2p_U_out"], out["c2p_V"], out["c2p_ent_out"], out["c2p_flag"] = \
        Uin, ent, Uout, Vc, eo, fl
    Vw = Vout[30:].copy()
    Vw[:, 4] = np.abs(Vw[:, 4]) + 1e-3
    sp = np.zeros((len(Vw), 3, 2))
    for i in range(len(Vw)):
        for d in range(3):
            a, b = np.zeros(1), np.zeros(1)
            lib.ref_wavspd(d + 1, gam, P(Vw[i].copy()), P(a), P(b))
            sp[i, d] = a[0], b[0]
    out["wav_V"], out["wav_S"] = Vw, sp
    # 2-D EMF solvers
    lib.ref_emf4_hll.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double] + [ctypes.c_void_p] * 4
    lib.ref_emf4_llf.argtypes = lib.ref_emf4_hll.argtypes
    lib.ref_emf3_llf.argtypes = [ctypes.c_void_p] * 4
    m = 600
    S = np.empty((m, 4))
    S[:, 0] = rng.uniform(-1, 3, m)            # S_R
    S[:, 1] = S[:, 0] - rng.uniform(0.1, 4, m) # S_L < S_R
    S[:, 2] = rng.uniform(-1, 3, m)            # S_U
    S[:, 3] = S[:, 2] - rng.uniform(0.1, 4, m) # S_D < S_U
    th = rng.uniform(0.6, 2.5, m)
    ez = rng.normal(0, 1, (m, 4))
    bn = rng.normal(0, 1, (m, 2))
    by = rng.normal(0, 1, (m, 2))
    o4 = np.zeros((m, 3))
    l4 = np.zeros((m, 3))
    for i in range(m):
        r = np.zeros(3)
        lib.ref_emf4_hll(P(S[i].copy()), np.cos(th[i]), np.sin(th[i]), P(ez[i].copy()), P(bn[i].copy()),
                         P(by[i].copy()), P(r))
        o4[i] = r
        r = np.zeros(3)
        lib.ref_emf4_llf(P(S[i].copy()), np.cos(th[i]), np.sin(th[i]), P(ez[i].copy()), P(bn[i].copy()),
                         P(by[i].copy()), P(r))
        l4[i] = r
    out["emf4_S"], out["emf4_theta"], out["emf4_ez"], out["emf4_bn"], out["emf4_by"] = S, th, ez, bn, by
    out["emf4_hll_out"], out["emf4_llf_out"] = o4, l4
    s6 = rng.uniform(-3, 3, (m, 6))
    e3 = rng.normal(0, 1, (m, 3))
    b3 = rng.normal(0, 1, (m, 3))
    o3 = np.zeros((m, 3))
    for i in range(m):
        r = np.zeros(3)
        lib.ref_emf3_llf(P(s6[i].copy()), P(e3[i].copy()), P(b3[i].copy()), P(r))
        o3[i] = r
    out["emf3_S6"], out["emf3_ez"], out["emf3_bn"], out["emf3_out"] = s6, e3, b3, o3
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)

    # quadpoints golden
    blocks = {}
    cur = None
    for line in open(REF_OUT):
        s = line.strip()
        if not s:
            continue
        if s.startswith("vert_coords_patch"):
            cur = "vert"
            blocks[cur] = []
        elif "_quadpts_patch" in s:
            cur = s.split(",")[0] + "_" + re.search(r"face (\d)", s).group(1)
            blocks[cur] = []
        elif "=" in s and cur == "vert":
            idx = [int(t) for t in re.findall(r"-?\d+", s.split("=")[0])]
            blocks[cur].append(idx + [float(t) for t in s.split("=")[1].split()])
        elif s.startswith("Point"):
            blocks[cur].append([float(t) for t in s.split(":")[1].split()])
    meta = {"source": "_T/_200mesh/test_quadpoints.out (written by _T/_200mesh/004test_quadpoints.f)",
            "setup": "sector 1, one zone per patch edge, NGhost 0, radii 10 and 11"}
    json.dump({"meta": meta, **blocks}, open(os.path.join(HERE, "quadpoints.json"), "w"), indent=1)
    print("wrote", sorted(out)[:4], "...", "and quadpoints.json with", sorted(blocks))


if __name__ == "__main__":
    sys.exit(main())
