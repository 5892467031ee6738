"""Third-order zone reconstruction (LIMIT_ALL_SLOPES_P12, 421limit_all_slopes_p12.f:9)
on the CPU: the oracle's restatement against the reference's own routines
built at NUMORDER 3 (oracle/_ref/libcsref3.so), and against closed forms.

The oracle restates the whole routine: the P1 WENO gradients (7.a) left
unflattened, the WENO-AO hybridisation of 7.b with the 18-zone quadratic
least-squares stencil (RECONSTRUCTION_COEFS, 251...f:3; MATINV,
121matinv.f:8) and the flattener, the positivity limiter with the quadratic
modes (7.c, 440...f:594-600), the face-point B of 8.a-c with them
(421...f:4817-4879), the 30-unknown constrained least squares of 8.d
(CMat_Row_Coefs_P2, 253...f:7; MATSOLVE), the zone E and its WENO-AO with
the quadratic interpolant of 9 (INTERPOLATION_COEFS, 252...f:5), and the
diagnostics of 10; every output array is compared."""
import ctypes
import os

import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

GEO_KEYS = ("v_pts", "v_wts", "v_tan", "v_vol", "v_cen", "moments")


def _ref3_or_skip():
    if orc.rlib3() is None:
        pytest.skip("third-order reference build (oracle/_ref/libcsref3.so) not present")


def _cfg(N=6, NR=3):
    return cs.make_config(N, NR, 4, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, rmin=1.0, rmax=1.6)


def test_volume_geometry3_matches_reference():
    """27-point rule (240...f:134-147) and ten moments (250...f:140-267) of
    the oracle against the reference's NUMORDER 3 build: points exact, the
    rest to 1e-13 (the serendipity derivatives are factored differently)."""
    _ref3_or_skip()
    cfg = _cfg()
    geo = orc.quad_geometry(cfg)
    a = orc.volume_geometry3(cfg, geo)
    b = orc.volume_geometry3(cfg, geo, use_reference=True)
    assert np.array_equal(a["v_pts"], b["v_pts"])
    for k in GEO_KEYS[1:]:
        scale = np.abs(b[k]).max()
        assert np.abs(a[k] - b[k]).max() <= 1e-13 * scale, k


def test_volume_geometry3_closed_forms():
    """Weights sum to one per zone; the centroid is the weighted point mean;
    the first moments vanish; the 27-point volume agrees with the 8-point one
    to the rule's error; second moments are positive."""
    cfg = _cfg()
    geo = orc.quad_geometry(cfg)
    a = orc.volume_geometry3(cfg, geo)
    nzn = orc.geometry_dims(cfg)["zone"]
    act = geo["mask"].reshape(6, nzn) == 1
    w = a["v_wts"].reshape(6, 27, nzn)
    assert np.allclose(w.sum(1)[act], 1.0, rtol=0, atol=1e-14)
    pts = a["v_pts"].reshape(6, 27, 3, nzn)
    cen = a["v_cen"].reshape(6, 3, nzn)
    mean = (w[:, :, None, :] * pts).sum(1)
    assert np.abs(mean - cen)[np.broadcast_to(act[:, None], cen.shape)].max() < 1e-13
    mom = a["moments"].reshape(6, 10, nzn)
    assert np.abs(mom[:, :4][np.broadcast_to(act[:, None], mom[:, :4].shape)]).max() < 1e-12
    assert (mom[:, 4:7][np.broadcast_to(act[:, None], mom[:, 4:7].shape)] > 0).all()
    v2 = geo["v_vol"].reshape(6, nzn)[act]
    v3 = a["v_vol"].reshape(6, nzn)[act]
    assert np.abs(v3 / v2 - 1.0).max() < 1e-4


def test_reconstruction_row_and_matinv_bit_exact_with_reference():
    _ref3_or_skip()
    L, R = orc.olib(), orc.rlib3()
    dbl, vp = ctypes.c_double, ctypes.c_void_p
    L.or_reconstruction_row3.argtypes = [dbl] * 5 + [vp] * 3
    R.ref3_reconstruction_row.argtypes = [dbl] * 5 + [vp] * 3
    L.or_matinv.argtypes = [ctypes.c_int, vp, vp]
    R.ref3_matinv.argtypes = [ctypes.c_int, vp, vp]
    rng = np.random.default_rng(5)
    for _ in range(200):
        x, y, z = rng.normal(size=3)
        l0, lj = rng.uniform(0.5, 2.0, size=2)
        c0, cj = rng.normal(size=10), rng.normal(size=10)
        a, b = np.zeros(9), np.zeros(9)
        L.or_reconstruction_row3(x, y, z, l0, lj, orc.P(c0), orc.P(cj), orc.P(a))
        R.ref3_reconstruction_row(x, y, z, l0, lj, orc.P(c0), orc.P(cj), orc.P(b))
        assert np.array_equal(a, b)
    for n in (3, 9, 18):
        for _ in range(20):
            M = rng.normal(size=(n, n)) + n * np.eye(n) * rng.uniform(0.2, 2.0)
            a, b = np.zeros((n, n)), np.zeros((n, n))
            m1 = M.copy()           # row-major for the oracle
            m2 = np.asfortranarray(M.copy())  # column-major for the reference
            L.or_matinv(n, orc.P(m1), orc.P(a))
            R.ref3_matinv(n, m2.ctypes.data_as(vp), b.ctypes.data_as(vp))
            assert np.array_equal(a, b.T), n  # b holds the column-major inverse
            assert np.abs(a @ M - np.eye(n)).max() < 1e-10


CASES = [(8, 3, 0, 3), (8, 3, 1, 3), (6, 4, 2, 7), (8, 3, 5, 11), (5, 2, 0, 11), (9, 3, 4, 13)]


def _hydro(A, nz):
    return A["V"].reshape(10, 8, nz)[:, :5]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d_nr%d_s%d_seed%d" % c)
def test_p12_zones_bit_exact_with_reference_smooth(case):
    """Smooth flow: the oracle's routine equals the reference's on every
    output array."""
    _ref3_or_skip()
    N, NR, s, seed = case
    prob, A0 = orc.p12_problem(N, NR, 4, sector=s, seed=seed, shock=False)
    nz = orc.p1_dims(N, NR, 4)["zone"]
    A = orc.p12_oracle(prob, orc.p1_copy(A0))
    B = orc.p12_reference(prob, orc.p1_copy(A0))
    # every output array of the routine: V, U, entropy, B_zc, E_zc, the
    # face-point B, the flattener and work arrays, the min/max
    bad = [k for k in A if not np.array_equal(A[k], B[k])]
    assert not bad, bad
    # the quadratic modes are there (not all zero)
    assert np.abs(_hydro(A, nz)[4:]).max() > 1e-4


@pytest.mark.parametrize("case", [(0, 3, 0.0), (3, 3, 0.0), (2, 5, 0.6), (5, 7, 1.5)],
                         ids=lambda c: "s%d_seed%d_noise%g" % c)
def test_p12_zones_with_flattener_and_positivity_bit_exact(case):
    """A compressive front (the flattener acts: the linear weights move to
    0.6 / 0.4 and every mode is scaled) and log-normal noise (the positivity
    limiter reduces slopes by 0.1^(n/15) or removes them, and its loop order
    matters): bit-exact with the reference on every output array."""
    _ref3_or_skip()
    sector, seed, noise = case
    N, NR = 8, 3
    prob, A0 = orc.p12_problem(N, NR, 4, sector=sector, seed=seed, shock=True, noise=noise)
    nz = orc.p1_dims(N, NR, 4)["zone"]
    A = orc.p12_oracle(prob, orc.p1_copy(A0))
    st = np.zeros(2, np.int64)
    orc.olib().or_p1_positivity_stats(orc.P(st))
    B = orc.p12_reference(prob, orc.p1_copy(A0))
    bad = [k for k in A if not np.array_equal(A[k], B[k])]
    assert not bad, bad
    assert (A["flatten"] > 0).sum() > 100
    assert st[0] > 5, "the positivity limiter must act"
    if noise > 1.0:
        assert st[1] > 5, "and remove slopes"


def test_p12_zones_bit_exact_with_nghost5():
    """The same routine on a patch with NGhost 5 (front and noise): the
    oracle's layouts and loop ranges follow the ghost width as the
    reference's do, bit-exact on every output array."""
    _ref3_or_skip()
    prob, A0 = orc.p12_problem(8, 3, 5, sector=3, seed=9, shock=True, noise=0.6)
    A = orc.p12_oracle(prob, orc.p1_copy(A0))
    B = orc.p12_reference(prob, orc.p1_copy(A0))
    bad = [k for k in A if not np.array_equal(A[k], B[k])]
    assert not bad, bad


def _quad_problem(N=8, NR=4, sector=1):
    prob, A = orc.p12_problem(N, NR, 4, sector=sector, seed=1)
    cfg = _cfg(N, NR)
    geo = orc.quad_geometry(cfg)
    g3 = orc.volume_geometry3(cfg, geo)
    gd = orc.geometry_dims(cfg)
    nz = gd["zone"]
    pts = g3["v_pts"][sector * nz * 81:(sector + 1) * nz * 81].reshape(27, 3, nz)
    wts = g3["v_wts"][sector * nz * 27:(sector + 1) * nz * 27].reshape(27, nz)
    return prob, A, pts, wts, nz


def test_p12_quadratic_fit_is_exact():
    """Closed form: zone means of a quadratic q (the 27-point rule's averages)
    are fitted exactly by the 18-zone stencil: X = (l0 grad q, l0^2 q_xx / 2,
    q_yy / 2, q_zz / 2, l0^2 q_xy, q_yz, q_xz) at the zone's centroid.  A wrong
    moment offset, scale ratio or mode order in the rows fails this."""
    N, NR = 8, 4
    prob, A, pts, wts, nz = _quad_problem(N, NR)
    g0 = np.array([0.3, -0.7, 0.45])
    H = np.array([[1.1, 0.4, -0.3], [0.4, -0.8, 0.25], [-0.3, 0.25, 0.6]])
    x0 = np.array([0.2, 0.1, 1.1])
    q = lambda x: 1.5 + g0 @ (x - x0[:, None]) + 0.5 * np.einsum("ip,ij,jp->p", x - x0[:, None], H, x - x0[:, None])
    Vm = (wts * np.stack([q(pts[p]) for p in range(27)])).sum(0)
    A["V"][:nz] = Vm  # component 0, mode 1: the zone means
    L = orc.olib()
    L.or_ao_fit_p12.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 4 + [ctypes.c_void_p]
    s = orc.p1_struct(prob, A)
    cen = A["v_centroid"].reshape(3, nz)
    d = orc.p1_dims(N, NR, 4)
    g = 4
    worst = 0.0
    for (i, j, k) in [(1, 1, 1), (4, 5, 2), (N, N, NR), (-1, 3, 0), (3, N + 2, NR + 2), (0, 0, 1)]:
        z = (i - 1 + g) + (N + 2 * g) * ((j - 1 + g) + (N + 2 * g) * (k - 1 + g))
        if A["mask"][z] != 1:
            continue
        X = np.zeros(9)
        assert L.or_ao_fit_p12(ctypes.byref(s), orc.P(A["moments"]), i, j, k, 0, orc.P(X)) == 0
        l0 = max(A["scale"][z], prob["smallnum"])
        c = cen[:, z]
        grad = g0 + H @ (c - x0)
        want = np.concatenate([l0 * grad, 0.5 * l0 ** 2 * np.diag(H), l0 ** 2 * np.array([H[0, 1], H[1, 2], H[0, 2]])])
        worst = max(worst, np.abs(X - want).max() / np.abs(want).max())
    assert worst < 1e-9, worst


def test_p12_linear_field_is_exact():
    """A linear primitive field: the WENO gradients and the quadratic fit are
    both exact, so 7.b returns the gradient (times 1 - flattener) and zero
    quadratic modes whatever the nonlinear weights."""
    N, NR = 8, 3
    prob, A, pts, wts, nz = _quad_problem(N, NR)
    cen = A["v_centroid"].reshape(3, nz)
    gvec = np.array([0.4, -0.2, 0.3])
    U = A["U"].reshape(10, 8, nz)
    U[0, 0] = 2.0 + gvec @ cen   # density linear in the centroid; at rest, uniform pressure, no B
    U[0, 1:4] = 0.0
    U[0, 4] = 1.0 / (prob["gamma"] - 1.0)
    U[0, 5:8] = 0.0
    A["ent"][:nz] = 1.0
    orc.p12_oracle(prob, A)
    V = A["V"].reshape(10, 8, nz)
    d = orc.p1_dims(N, NR, 4)
    g = 4
    worst_g, worst_q = 0.0, 0.0
    for k in range(1, NR + 1):
        for j in range(1, N + 1):
            for i in range(1, N + 1):
                z = (i - 1 + g) + (N + 2 * g) * ((j - 1 + g) + (N + 2 * g) * (k - 1 + g))
                l0 = max(A["scale"][z], prob["smallnum"])
                want = l0 * gvec * (1.0 - A["flatten"][z])
                worst_g = max(worst_g, np.abs(V[1:4, 0, z] - want).max() / np.abs(l0 * gvec).max())
                worst_q = max(worst_q, np.abs(V[4:, 0, z]).max() / np.abs(l0 * gvec).max())
    assert worst_g < 1e-10 and worst_q < 1e-10, (worst_g, worst_q)


# zone offsets of the 18-zone stencil: direction in the stencil table (-1 for
# the zone's own column) and the radial offset, as INTERPOLATION_COEFS' rows
# are filled at 421limit_all_slopes_p12.f:8899-8935
_AO_DIR = (0, 1, 2, 3, -1, -1, 8, 9, 10, 11, 0, 1, 0, 1, 2, 3, 2, 3)
_AO_DK = (0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1)


@pytest.mark.parametrize("N", [8, 12, 24, 192])
def test_step9c_levels_follow_the_serial_loop_order(N):
    """Step 9.c overwrites each zone's point value of E while later zones of
    the loop read it; the GPU runs it level by level, level
    (i+1) + 3 (j+1) + 9 (k+1) over the box -1..N+2.  On the stencil table of
    the cubed-sphere patch (corner strips included) every active neighbour
    that the serial (k, j, i) loop visits earlier has a lower level and every
    later one a higher level, and no entry reaches beyond two zones in i or j
    (the bound that the GPU's own check enforces); at the corner strips the
    plainer level i + 2j + 4k breaks this order."""
    g = 4
    nz = N + 2 * g
    st = np.zeros(nz * nz * 24, np.int32)
    L = orc.olib()
    L.or_stencil_indexing.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.or_stencil_indexing(N, g, st.ctypes.data)
    st = st.reshape(24, nz, nz)
    prob, A = orc.p1_problem(N, 2, g, sector=0, seed=1)
    mask = A["mask"].reshape(-1, nz, nz)[g]

    def act(i, j):
        return -1 <= i <= N + 2 and -1 <= j <= N + 2 and mask[j - 1 + g, i - 1 + g] == 1

    def violations(wj, wk):
        bad, reach = 0, 0
        for j in range(-1, N + 3):
            for i in range(-1, N + 3):
                if not act(i, j):
                    continue
                for r in range(18):
                    a, b = i, j
                    if _AO_DIR[r] >= 0:
                        a = int(st[2 * _AO_DIR[r], j - 1 + g, i - 1 + g])
                        b = int(st[2 * _AO_DIR[r] + 1, j - 1 + g, i - 1 + g])
                    if not act(a, b):
                        continue
                    dk = _AO_DK[r]
                    reach = max(reach, abs(a - i), abs(b - j))
                    earlier = (dk, b - j, a - i) < (0, 0, 0)
                    dl = (a - i) + wj * (b - j) + wk * dk
                    bad += int(dl >= 0 if earlier else dl <= 0)
        return bad, reach

    bad, reach = violations(3, 9)
    assert bad == 0 and reach == 2
    assert violations(2, 4)[0] > 0


def test_p12_oracle_matches_reference_digests():
    """Without the reference build: the oracle's LIMIT_ALL_SLOPES_P12 leaves,
    bit for bit, the arrays the reference's own routine left on the same
    inputs (SHA-256 of every array, tests/golden/p12_digests.json, written by
    tests/golden/make_golden_p12.py from the reference built at NUMORDER 3).
    The inputs' digests are checked first, so a change in the problem
    generator is told apart from a change in the routine."""
    import hashlib
    import json
    path = os.path.join(os.path.dirname(__file__), "golden", "p12_digests.json")
    cases = json.load(open(path))
    assert len(cases) == 3

    def dig(a):
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()

    for c in cases:
        N, NR, g, s, seed, shock, noise = c["case"]
        prob, A0 = orc.p12_problem(N, NR, g, sector=s, seed=seed, shock=shock, noise=noise)
        bad_in = [k for k in sorted(A0) if dig(A0[k]) != c["inputs"].get(k)]
        assert not bad_in and set(A0) == set(c["inputs"]), ("inputs differ", c["case"], bad_in)
        A = orc.p12_oracle(prob, orc.p1_copy(A0))
        bad = [k for k in sorted(A) if dig(A[k]) != c["reference"].get(k)]
        assert not bad and set(A) == set(c["reference"]), (c["case"], bad)
