"""GPU parity of the third-order entry points with the oracle, which is
bit-exact with the reference built at NUMORDER 3 (tests/test_cpu_p12.py):
cs_build_volume_p12 (COORDS_TO_WT_QUADPTS_VOLUME_ISO2 with 27 points and ten
basis moments) and cs_limit_all_slopes_p12 (LIMIT_ALL_SLOPES_P12,
421limit_all_slopes_p12.f:9).

Tolerance: the HIP path fuses multiply-adds and solves the 9 x 9 normal
equations of the quadratic stencil by Cholesky where the reference inverts
them by LU, and the 36 x 36 system of step 8.d by elimination across a
wave's lanes where the reference uses LU, so the geometry agrees to 1e-13 and each (component, mode) row of
V, B*_zc, E*_zc and the face-point B to 1e-9 of its largest magnitude; the flattener, the entropy and
the integer arrays to 1e-10 / exactly.  The positivity limiter's discrete
decisions (how many reductions, removal) must be the same: a different
decision would move a mode by 14 % or more, far outside the tolerance."""
import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_build_volume_p12_matches_oracle():
    cfg = cs.make_config(8, 4, 4, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, rmin=1.0, rmax=1.6)
    print("case volume 8x4", flush=True)
    got = cs.build_volume_p12(cfg)
    geo = orc.quad_geometry(cfg)
    want = orc.volume_geometry3(cfg, geo)
    for k in ("v_pts", "v_wts", "v_vol", "v_cen", "moments"):
        scale = np.abs(want[k]).max()
        err = np.abs(got[k] - want[k]).max()
        assert err <= 1e-13 * scale, (k, err, scale)


CASES = [
    # N, NR, sector, seed, shock, noise
    (8, 3, 0, 3, False, 0.0),
    (8, 3, 1, 3, True, 0.0),
    (12, 5, 4, 2, True, 0.6),
    (16, 6, 2, 5, False, 1.5),
    (10, 4, 5, 7, True, 1.0),
    (8, 3, 3, 9, True, 0.6, 5),  # NGhost 5
    (5, 2, 0, 11, True, 0.8),  # a patch barely wider than its ghosts
    (9, 3, 4, 13, False, 0.3),  # odd N
]


def _run_gpu(prob, A):
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A[k]).to(dev) for k in cs.P1_ARRAYS}
    mom = torch.from_numpy(A["moments"]).to(dev)
    vmin, vmax = cs.limit_all_slopes_p12(prob["N"], prob["NR"], prob["g"], prob, T, mom)
    torch.cuda.synchronize()
    out = {k: T[k].cpu().numpy() for k in cs.P1_ARRAYS}
    out["V_min_local"], out["V_max_local"] = vmin, vmax
    return out


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d_nr%d_s%d_seed%d_%s_noise%g_g%d" % (
    c[0], c[1], c[2], c[3], "front" if c[4] else "smooth", c[5], c[6] if len(c) > 6 else 4))
def test_p12_zones_match_oracle(case):
    N, NR, s, seed, shock, noise = case[:6]
    g = case[6] if len(case) > 6 else 4
    print(f"case p12 zones {case}", flush=True)
    prob, A0 = orc.p12_problem(N, NR, g, sector=s, seed=seed, shock=shock, noise=noise)
    O = orc.p12_oracle(prob, orc.p1_copy(A0))
    G = _run_gpu(prob, orc.p1_copy(A0))
    nz = orc.p1_dims(N, NR, g)["zone"]
    bad = []
    for k in cs.P1_ARRAYS + ("V_min_local", "V_max_local"):
        g, o = G[k], O[k]
        if o.dtype.kind != "f":
            if not np.array_equal(g, o):
                bad.append((k, "int"))
            continue
        rows = 80 if k in ("U", "V") else (4 if k == "ent" else (10 if k.endswith("_zc") else
                                                                   (5 if k.endswith("_q") else
                                                                    (9 if k.startswith("V_m") else 1))))
        rtol = 1e-9 if (k == "V" or k.endswith("_zc") or k.endswith("_q")) else 1e-10
        for r, (gr, orow) in enumerate(zip(g.reshape(rows, -1), o.reshape(rows, -1))):
            scale = max(np.abs(orow).max(), 1e-300)
            err = np.abs(gr - orow).max()
            if not err <= rtol * scale:
                bad.append((k, r, float(err), float(scale)))
    assert not bad, bad[:10]
    # the quadratic modes were formed, and the flattener acted in the front cases
    V = G["V"].reshape(10, 8, nz)
    assert np.abs(V[4:]).max() > 1e-5
    if shock:
        assert (G["flatten"] > 0).sum() > 50


def test_p12_rejects_a_stencil_that_breaks_the_step9c_order():
    """Zone (1, 1)'s right neighbour moved to (1, 4), three rows up (an index
    inside the patch, so every kernel still reads in bounds): its stencil zone
    below the plane is then on the same level as the zone, and the call
    reports that it cannot keep the serial loop order."""
    N, NR, g = 8, 3, 4
    print("case p12 stencil order", flush=True)
    prob, A0 = orc.p12_problem(N, NR, g, sector=0, seed=3)
    A = orc.p1_copy(A0)
    nz = N + 2 * g
    s = (1 - 1 + g) + nz * (1 - 1 + g)
    assert A["stencil"][s] == 2 and A["stencil"][s + nz * nz] == 1  # direction 0: (2, 1)
    A["stencil"][s + nz * nz] = 4
    with pytest.raises(cs.CsError, match="level order"):
        _run_gpu(prob, A)


def test_p12_timed_matches_untimed():
    """cs_limit_all_slopes_p12_timed: the same arrays as the plain call, and
    eight stage times in milliseconds."""
    N, NR = 10, 4
    print("case p12 timed", flush=True)
    prob, A0 = orc.p12_problem(N, NR, 4, sector=2, seed=4, shock=True, noise=0.5)
    G = _run_gpu(prob, orc.p1_copy(A0))
    dev = torch.device("cuda:0")
    A = orc.p1_copy(A0)
    T = {k: torch.from_numpy(A[k]).to(dev) for k in cs.P1_ARRAYS}
    mom = torch.from_numpy(A["moments"]).to(dev)
    vmin, vmax, ms = cs.limit_all_slopes_p12(prob["N"], prob["NR"], prob["g"], prob, T, mom, timed=True)
    torch.cuda.synchronize()
    assert len(ms) == 8 and all(t >= 0.0 for t in ms) and sum(ms) > 0.0
    for k in cs.P1_ARRAYS:
        assert np.array_equal(T[k].cpu().numpy(), G[k]), k
    assert np.array_equal(vmin, G["V_min_local"]) and np.array_equal(vmax, G["V_max_local"])


def test_p12_linear_field_is_exact_at_48x12():
    """A size-independent property, at a size the oracle would take minutes
    over: a linear density at rest (uniform pressure, no B) is reconstructed
    exactly by both the WENO gradients and the 18-zone quadratic fit, so every
    owned zone gets the gradient (times the scale length and 1 - flattener)
    and no quadratic modes, whatever the nonlinear weights; B and E stay
    zero."""
    N, NR, g = 48, 12, 4
    print("case p12 linear field 48x12", flush=True)
    prob, A = orc.p12_problem(N, NR, g, sector=1, seed=1)
    nz = orc.p1_dims(N, NR, g)["zone"]
    cen = A["v_centroid"].reshape(3, nz)
    gvec = np.array([0.4, -0.2, 0.3])
    U = A["U"].reshape(10, 8, nz)
    U[:] = 0.0
    U[0, 0] = 2.0 + gvec @ cen
    U[0, 4] = 1.0 / (prob["gamma"] - 1.0)
    for k in ("Bxf", "Byf", "Bzf"):
        A[k][:] = 0.0
    A["ent"][:] = 0.0
    A["ent"][:nz] = 1.0
    G = _run_gpu(prob, A)
    V = G["V"].reshape(10, 8, nz)
    nzz = N + 2 * g
    ii = np.arange(1, N + 1)
    kk = np.arange(1, NR + 1)
    own = ((ii[None, None, :] - 1 + g) + nzz * ((ii[None, :, None] - 1 + g) + nzz * (kk[:, None, None] - 1 + g))).ravel()
    l0 = np.maximum(G["scale"][own], prob["smallnum"])
    want = l0[None, :] * gvec[:, None] * (1.0 - G["flatten"][own])[None, :]
    ref = np.abs(l0[None, :] * gvec[:, None]).max()
    err_g = np.abs(V[1:4, 0, own] - want).max() / ref
    err_q = np.abs(V[4:, 0, own]).max() / ref
    assert err_g < 1e-9 and err_q < 1e-9, (err_g, err_q)
    for k in ("B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc"):
        assert np.abs(G[k].reshape(10, nz)[:, own]).max() == 0.0, k
