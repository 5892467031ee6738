"""GPU parity of cs_limit_all_slopes_p1 (LIMIT_ALL_SLOPES_P1,
410limit_all_slopes_p1.f:9) with the oracle, which is bit-exact with the
reference's own routine (tests/test_cpu_p1.py).

Tolerance: the HIP path evaluates the same FP64 formulas, but lets the
compiler fuse multiply-adds and solves the constrained least squares through
a Cholesky / Schur complement instead of the reference's LU, so each array
agrees to 1e-10 of its largest magnitude; integer arrays, the flattener's
inputs and every discrete decision (flattening, positivity reductions,
complete removal) agree exactly."""
import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = [
    # N, NR, sector, seed, shock, noise
    (8, 3, 0, 0, True, 0.0),
    (8, 4, 3, 1, True, 0.3),
    (12, 5, 1, 2, True, 1.0),
    (10, 4, 5, 3, False, 2.0),
    (24, 8, 2, 4, True, 0.5),
]


def _run_gpu(prob, A):
    dev = torch.device("cuda:0")
    T = {}
    for k in cs.P1_ARRAYS:
        T[k] = torch.from_numpy(A[k]).to(dev)
    vmin, vmax = cs.limit_all_slopes_p1(prob["N"], prob["NR"], prob["g"], prob, T)
    torch.cuda.synchronize()
    out = {k: T[k].cpu().numpy() for k in cs.P1_ARRAYS}
    out["V_min_local"] = vmin
    out["V_max_local"] = vmax
    return out


def _compare(G, O, rtol=1e-10, rows=None):
    """Each array to rtol of its largest magnitude; with rows = {name: n}, the
    array is split into n equal rows (components, modes) and each row is
    scaled by its own largest magnitude."""
    bad = []
    rows = rows or {}
    for k in O:
        g, o = G[k], O[k]
        if o.dtype.kind != "f":
            if not np.array_equal(g, o):
                bad.append((k, "int mismatch", int((g != o).sum())))
            continue
        n = rows.get(k, 1)
        for r, (gr, orow) in enumerate(zip(g.reshape(n, -1), o.reshape(n, -1))):
            scale = max(np.abs(orow).max(), 1e-300)
            err = np.abs(gr - orow).max()
            if not err <= rtol * scale:
                bad.append((k, r, float(err), float(scale), int((np.abs(gr - orow) > rtol * scale).sum())))
    return bad


@pytest.mark.parametrize("case", CASES, ids=lambda c: "N%d_nr%d_s%d_seed%d_%s_noise%g" % (
    c[0], c[1], c[2], c[3], "front" if c[4] else "smooth", c[5]))
def test_p1_gpu_matches_oracle(case):
    N, NR, s, seed, shock, noise = case
    print(f"\np1 case N={N} NR={NR} sector={s} seed={seed} noise={noise}", flush=True)
    prob, A0 = orc.p1_problem(N, NR, 4, sector=s, seed=seed, shock=shock, noise=noise)
    O = orc.p1_oracle(prob, orc.p1_copy(A0))
    G = _run_gpu(prob, orc.p1_copy(A0))
    bad = _compare(G, {k: O[k] for k in G})
    assert not bad, bad
    # discrete decisions agree exactly
    nz = orc.p1_dims(N, NR, 4)["zone"]
    np.testing.assert_array_equal(G["flatten"] > 0, O["flatten"] > 0)
    np.testing.assert_array_equal(G["ent"].reshape(4, nz)[0] == 0, O["ent"].reshape(4, nz)[0] == 0)
    vg, vo = G["V"].reshape(4, 8, nz), O["V"].reshape(4, 8, nz)
    np.testing.assert_array_equal(np.abs(vg[1:, :5]).sum((0, 1)) == 0, np.abs(vo[1:, :5]).sum((0, 1)) == 0)


def test_p1_gpu_rejects_bad_sizes():
    prob, A = orc.p1_problem(8, 3, 4)
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A[k]).to(dev) for k in cs.P1_ARRAYS}
    T["V"] = T["V"][:-1].contiguous()
    with pytest.raises(cs.CsError):
        cs.limit_all_slopes_p1(8, 3, 4, prob, T)
    T = {k: torch.from_numpy(A[k]).to(dev) for k in cs.P1_ARRAYS}
    with pytest.raises(cs.CsError):
        cs.limit_all_slopes_p1(8, 3, 3, prob, T)  # nghost < 4: sizes and the entry point refuse


FLUX_CASES = [
    (8, 3, 0, 0, True, 0.0),
    (12, 5, 1, 2, True, 1.0),
    (10, 4, 5, 3, False, 0.5),
    (24, 8, 2, 4, True, 0.3),
]


@pytest.mark.parametrize("case", FLUX_CASES, ids=lambda c: "N%d_nr%d_s%d_seed%d_%s_noise%g" % (
    c[0], c[1], c[2], c[3], "front" if c[4] else "smooth", c[5]))
def test_facial_fluxes_gpu_matches_oracle(case):
    """cs_make_facial_fluxes_p1 against or_make_facial_fluxes_p1 on a patch the
    oracle's LIMIT_ALL_SLOPES_P1 reconstructed: every output to 1e-11 of its
    largest magnitude (the wave speeds use a reciprocal, the rest the same
    formulas), the tri-corner re-pointed shares included."""
    N, NR, s, seed, shock, noise = case
    print(f"\nflux case N={N} NR={NR} sector={s} seed={seed} noise={noise}", flush=True)
    prob, A0 = orc.p1_flux_problem(N, NR, 4, sector=s, seed=seed, shock=shock, noise=noise)
    O = orc.p1f_oracle(prob, orc.p1_copy(A0))
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A0[k]).to(dev) for k in cs.P1F_IN + cs.P1F_OUT}
    cs.make_facial_fluxes_p1(N, NR, 4, prob, T)
    torch.cuda.synchronize()
    G = {k: T[k].cpu().numpy() for k in cs.P1F_OUT}
    bad = _compare(G, {k: O[k] for k in cs.P1F_OUT}, rtol=1e-11)
    assert not bad, bad


@pytest.mark.parametrize("case", [(8, 3, 0, 0, 0.0), (12, 5, 1, 2, 1.0), (24, 8, 2, 4, 0.3)],
                         ids=lambda c: "N%d_nr%d_s%d_seed%d_noise%g" % c)
def test_update_gpu_matches_oracle(case):
    """cs_update_imex_rk_step_p1 against or_update_imex_rk_step_p1 on a patch
    after the oracle's reconstruction and facial fluxes: U, entropy and face B
    to 1e-12 of their largest magnitude (acos and the entropy's power round
    within an ulp of the host's), the flags exactly."""
    N, NR, s, seed, noise = case
    print(f"\nupdate case N={N} NR={NR} sector={s} seed={seed}", flush=True)
    prob, A0 = orc.p1_update_problem(N, NR, 4, sector=s, seed=seed, dt=2e-3, noise=noise)
    O = orc.p1u_oracle(prob, orc.p1_copy(A0))
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A0[k]).to(dev) for k in cs.P1U_ARRAYS}
    cs.update_imex_rk_step_p1(N, NR, 4, prob, T)
    torch.cuda.synchronize()
    G = {k: T[k].cpu().numpy() for k in ("U", "ent", "Bxf", "Byf", "Bzf", "flag_entprs")}
    bad = _compare(G, {k: O[k] for k in G}, rtol=1e-12)
    assert not bad, bad
    np.testing.assert_array_equal(G["flag_entprs"], O["flag_entprs"])


CLEANUP_CASES = [(8, 3, 0, 0, 0.0, True), (12, 5, 1, 2, 1.0, True), (10, 4, 5, 3, 0.5, False),
                 (24, 8, 2, 4, 0.3, True)]


@pytest.mark.parametrize("case", CLEANUP_CASES, ids=lambda c: "N%d_nr%d_s%d_seed%d_noise%g_floors%d" % c)
def test_cleanup_and_timestep_gpu_match_oracle(case):
    """cs_cleanup_after_update_p1 against or_cleanup_after_update_p1 (bit-exact
    with the reference's CLEANUP_AFTER_UPDATE) on a patch after the oracle's
    reconstruction, fluxes and update, with floored zones: every array it
    writes to 1e-12 of the largest magnitude of its component and mode (the
    3 x 3 normal equations and CONS_TO_PRIM with fused multiply-adds), then cs_evaluate_timestep_p1 on the
    cleaned patch: the time step to 1e-13 relative, U, V and msonic to 1e-12."""
    N, NR, s, seed, noise, floors = case
    print(f"\ncleanup case N={N} NR={NR} sector={s} seed={seed} floors={floors}", flush=True)
    prob, A0 = orc.p1_cleanup_problem(N, NR, 4, sector=s, seed=seed, noise=noise, floors=floors)
    O = orc.p1c_oracle(prob, orc.p1_copy(A0))
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A0[k]).to(dev) for k in cs.P1_ARRAYS}
    cs.cleanup_after_update_p1(N, NR, 4, prob, T)
    torch.cuda.synchronize()
    keys = ("U", "ent", "V", "Bxf", "Byf", "Bzf", "B1_zc", "B2_zc", "B3_zc", "scale", "gamma_eos", "gamma_sound",
            "mean_mol_wt")
    G = {k: T[k].cpu().numpy() for k in keys}
    rows = dict(U=32, V=32, ent=4, B1_zc=4, B2_zc=4, B3_zc=4)
    bad = _compare(G, {k: O[k] for k in keys}, rtol=1e-12, rows=rows)
    assert not bad, bad
    # the time step from the oracle's cleaned state on both sides
    for dtcur in (1.0, 1e-18):
        A1 = orc.p1_copy(O)
        T = {k: torch.from_numpy(A1[k]).to(dev) for k in cs.P1_ARRAYS}
        dt_g = float(cs.evaluate_timestep_p1(N, NR, 4, prob, T, 0.4, dtcur).item())
        dt_o = orc.p1_timestep_oracle(prob, A1, 0.4, dtcur)
        assert abs(dt_g - dt_o) <= 1e-13 * dt_o, (dt_g, dt_o)
        G = {k: T[k].cpu().numpy() for k in ("U", "V", "msonic", "ent")}
        bad = _compare(G, {k: A1[k] for k in G}, rtol=1e-12, rows=rows)
        assert not bad, bad


def test_timestep_gpu_rejects_bad_arguments():
    prob, A = orc.p1_problem(8, 3, 4)
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A[k]).to(dev) for k in cs.P1_ARRAYS}
    with pytest.raises(cs.CsError):
        cs.evaluate_timestep_p1(8, 3, 4, prob, T, 0.0, 1.0)
    T["msonic"] = T["msonic"][:-1].contiguous()
    with pytest.raises(cs.CsError):
        cs.evaluate_timestep_p1(8, 3, 4, prob, T, 0.4, 1.0)


EDGE_CASES = [(8, 3, 0, 0, True, 0.0, (1, 1, 1, 1)), (12, 5, 1, 2, True, 1.0, (1, 1, 1, 1)),
              (10, 4, 5, 3, False, 0.5, (1, 1, 1, 1)), (24, 8, 2, 4, True, 0.3, (1, 0, 1, 0))]


@pytest.mark.parametrize("case", EDGE_CASES, ids=lambda c: "N%d_nr%d_s%d_seed%d_%s_noise%g_itri%s" % (
    c[0], c[1], c[2], c[3], "front" if c[4] else "smooth", c[5], "".join(map(str, c[6]))))
def test_edge_efields_gpu_match_oracle(case):
    """cs_make_edge_efields_p1 against or_make_edge_efields_p1 (bit-exact with
    the reference's MAKE_EDGE_EFIELDS_P1) on a patch after the oracle's
    reconstruction and facial fluxes, the edge arrays starting as junk: Efx,
    Efy, Efz to 1e-12 of their largest magnitude (the tangents normalised
    with fused multiply-adds), the zeros off the owned edges exactly."""
    N, NR, s, seed, shock, noise, itri = case
    print(f"\nedge case N={N} NR={NR} sector={s} seed={seed} itri={itri}", flush=True)
    prob, A0 = orc.p1_edge_problem(N, NR, 4, sector=s, seed=seed, shock=shock, noise=noise)
    O = orc.p1e_oracle(prob, orc.p1_copy(A0), itri)
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A0[k]).to(dev) for k in cs.P1E_IN + cs.P1E_OUT}
    cs.make_edge_efields_p1(N, NR, 4, T, itri)
    torch.cuda.synchronize()
    G = {k: T[k].cpu().numpy() for k in cs.P1E_OUT}
    bad = _compare(G, {k: O[k] for k in cs.P1E_OUT}, rtol=1e-12)
    assert not bad, bad
    for k in cs.P1E_OUT:
        np.testing.assert_array_equal(G[k] == 0, O[k] == 0)


@pytest.mark.parametrize("case", [(12, 5, 1, 3, 0.3), (16, 6, 4, 8, 1.0)], ids=lambda c: "N%d_nr%d_s%d_seed%d_noise%g" % c)
def test_p1_stage_gpu_matches_oracle(case):
    """The reference's per-patch stage on one patch, from the raw state:
    LIMIT_ALL_SLOPES_P1, MAKE_FACIAL_FLUXES, MAKE_EDGE_EFIELDS_P1,
    BUILD_NON_STIFF_SOURCES (point mass and rotation on),
    UPDATE_IMEX_RK_STEP_P1, CLEANUP_AFTER_UPDATE, EVALUATE_TIMESTEP, every
    routine on the GPU through the C-ABI against the same sequence of the
    oracle's routines (each bit-exact with the reference's or pinned by its
    invariants).  Rounding differences pass through the limiters, so the
    state after the stage is compared to 1e-9 of each component's and mode's
    largest magnitude, the new time step to 1e-9 relative."""
    N, NR, s, seed, noise = case
    print(f"\nstage case N={N} NR={NR} sector={s} seed={seed} noise={noise}", flush=True)
    prob, A0 = orc.p1_problem(N, NR, 4, sector=s, seed=seed, shock=True, noise=noise)
    d = orc.p1_dims(N, NR, 4)
    for k in cs.P1F_OUT:
        A0[k] = np.zeros((8 if k.startswith("Fluct") else 1) * d["fc"])
    A0["dU_dt"] = np.zeros(8 * d["zone"])
    A0["flag_entprs"] = np.zeros(d["zone"])
    A0["Source"] = np.zeros(8 * d["zone"])
    uprob = dict(prob, dtmove=2e-3)
    sprob = dict(prob, newtonsc=1.0, mstar=0.5, omega=(0.05, 0.0, 0.2))
    O = orc.p1_copy(A0)
    orc.p1_oracle(prob, O)
    orc.p1f_oracle(prob, O)
    orc.p1e_oracle(prob, O)
    orc.p1s_oracle(sprob, O)
    orc.p1u_oracle(uprob, O)
    orc.p1c_oracle(prob, O)
    dt_o = orc.p1_timestep_oracle(prob, O, 0.4, 1.0)
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(v).to(dev) for k, v in A0.items()}
    cs.limit_all_slopes_p1(N, NR, 4, prob, T)
    cs.make_facial_fluxes_p1(N, NR, 4, prob, T)
    cs.make_edge_efields_p1(N, NR, 4, T)
    cs.build_non_stiff_sources_p1(N, NR, 4, 1.0, 0.5, prob["smallnum"], sprob["omega"], T)
    cs.update_imex_rk_step_p1(N, NR, 4, uprob, T)
    cs.cleanup_after_update_p1(N, NR, 4, prob, T)
    dt_g = float(cs.evaluate_timestep_p1(N, NR, 4, prob, T, 0.4, 1.0).item())
    keys = ("U", "ent", "V", "Bxf", "Byf", "Bzf", "Efx", "Efy", "Efz", "Source", "flag_entprs")
    G = {k: T[k].cpu().numpy() for k in keys}
    bad = _compare(G, {k: O[k] for k in keys}, rtol=1e-9, rows=dict(U=32, V=32, ent=4, Source=8))
    assert not bad, bad
    assert abs(dt_g - dt_o) <= 1e-9 * dt_o, (dt_g, dt_o)
    np.testing.assert_array_equal(G["flag_entprs"], O["flag_entprs"])


@pytest.mark.parametrize("gm,omega", [((1.0, 0.8), (0.1, -0.2, 0.3)), ((2.0, 0.5), (0.0, 0.0, 0.0)),
                                      ((1.0, 0.0), (0.0, 0.4, -0.1)), ((1.0, 0.0), (0.0, 0.0, 0.0))])
def test_sources_gpu_match_oracle(gm, omega):
    """cs_build_non_stiff_sources_p1 against or_build_non_stiff_sources_p1 on a
    reconstructed patch (gravity and rotation, each alone, and neither):
    Source and dU_dt to 1e-13 of each component's largest magnitude, Source
    exactly zero where the oracle's is."""
    print(f"\nsources gm={gm} omega={omega}", flush=True)
    N, NR = 12, 5
    prob, A0 = orc.p1_source_problem(N, NR, 4, sector=3, seed=6, gm=gm, omega=omega)
    O = orc.p1s_oracle(prob, orc.p1_copy(A0))
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A0[k]).to(dev) for k in orc.P1S_ARRAYS}
    cs.build_non_stiff_sources_p1(N, NR, 4, gm[0], gm[1], prob["smallnum"], omega, T)
    torch.cuda.synchronize()
    G = {k: T[k].cpu().numpy() for k in ("Source", "dU_dt")}
    bad = _compare(G, {k: O[k] for k in G}, rtol=1e-13, rows=dict(Source=8, dU_dt=8))
    assert not bad, bad
    np.testing.assert_array_equal(G["Source"] == 0, O["Source"] == 0)


@pytest.mark.parametrize("N,NR,sector,seed,bc", [(12, 5, 0, 1, (2, 2)), (10, 4, 4, 2, (2, 17)),
                                                (9, 6, 2, 3, (17, 2))])
def test_pad_boundary_r_gpu_bit_exact(N, NR, sector, seed, bc):
    """cs_pad_boundary_r_p1 against or_pad_boundary_r_p1 (which matches the
    reference's PAD_BOUNDARY_R bit for bit): U with its four modes, Bxf, Byf,
    Bzf bit for bit, for both sides padded and each side alone."""
    print(f"\npad N={N} NR={NR} sector={sector} bc={bc}", flush=True)
    prob, A0 = orc.p1_pad_problem(N, NR, 4, sector=sector, seed=seed, inner=bc[0], outer=bc[1])
    O = orc.p1p_oracle(prob, orc.p1_copy(A0))
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A0[k]).to(dev) for k in orc.P1P_ARRAYS}
    cs.pad_boundary_r_p1(N, NR, 4, T)
    torch.cuda.synchronize()
    for k in ("U", "Bxf", "Byf", "Bzf"):
        np.testing.assert_array_equal(T[k].cpu().numpy(), O[k], err_msg=k)


def test_pad_boundary_r_gpu_rejects_bad_sizes():
    """A U of fewer than eight components, or a bcarr_r of the wrong size, is
    refused on the host."""
    prob, A0 = orc.p1_pad_problem(8, 3, 4)
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A0[k]).to(dev) for k in orc.P1P_ARRAYS}
    for k, bad in (("U", T["U"][:7 * T["mask"].numel()]), ("bcarr_r", T["bcarr_r"][:-1])):
        with pytest.raises(cs.CsError):
            cs.pad_boundary_r_p1(8, 3, 4, dict(T, **{k: bad}))


def test_p1_shell_steps_gpu_match_oracle():
    """Two Heun RK2 steps of the driver's loop over the six P1 patches of a
    shell (cubesphere_amd.P1Shell: cs_halo_exchange_p1, every P1 entry point
    and the defect correction cs_sync_shared_p1 through the C-ABI) against
    the same loop of the oracle's routines (oracle.p1_shell_step_oracle, the
    exchange being the oracle's geometric one, the same face and edge
    pairings): every patch's U, entropy and face B to 1e-9 of each
    component's and mode's largest magnitude, the next time step to 1e-9
    relative."""
    N, NR, g = 8, 4, 4
    print(f"\nshell N={N} NR={NR}", flush=True)
    cfg, probs, As = orc.p1_shell_problem(N, NR, g, seed=3)
    grid = orc.Grid(cfg)
    src = (1.0, 0.3, (0.0, 0.1, 0.2))
    dt = min(orc.p1_timestep_oracle(p, orc.p1_copy(A), 0.4, 1.0) for p, A in zip(probs, As))
    dev = torch.device("cuda:0")
    Ts = [{k: torch.from_numpy(v.copy()).to(dev) for k, v in A.items()} for A in As]
    shell = cs.P1Shell(cfg, Ts, probs[0], sources=src)
    pairings = cs.p1_shell_pairings(As, N, NR, g)
    try:
        for _ in range(2):
            orc.p1_shell_step_oracle(grid, probs, As, dt, sources=src, pairings=pairings)
            shell.step(dt)
        torch.cuda.synchronize()
        for s in range(6):
            G = {k: Ts[s][k].cpu().numpy() for k in ("U", "ent", "Bxf", "Byf", "Bzf")}
            bad = _compare(G, {k: As[s][k] for k in G}, rtol=1e-9, rows=dict(U=32, ent=4))
            assert not bad, (s, bad)
        dt_g = shell.timestep(0.4, dt)
        dt_o = min(orc.p1_timestep_oracle(p, A, 0.4, dt) for p, A in zip(probs, As))
        assert abs(dt_g - dt_o) <= 1e-9 * dt_o, (dt_g, dt_o)
    finally:
        shell.close()


def _shell_after_fluxes(N, NR, g, seed):
    """six patches after the oracle's exchange, reconstruction, facial fluxes
    and edge fields"""
    cfg, probs, As = orc.p1_shell_problem(N, NR, g, seed=seed)
    grid = orc.Grid(cfg)
    orc.p1_shell_exchange_oracle(grid, As)
    for p, A in zip(probs, As):
        orc.p1_oracle(p, A)
        orc.p1f_oracle(p, A)
        orc.p1e_oracle(p, A)
    return cfg, probs, As, grid


def test_p1_halo_exchange_gpu_bit_exact():
    """cs_halo_exchange_p1 over six P1 patches against the oracle's geometric
    exchange of the same mode-1 arrays: U, entropy and face B bit for bit
    (modes 2-4 untouched)."""
    N, NR, g = 8, 4, 4
    cfg, probs, As = orc.p1_shell_problem(N, NR, g, seed=4)
    rng = np.random.default_rng(9)
    for A in As:  # junk in the ghosts, so the exchange has something to overwrite
        for k in ("U", "ent", "Bxf", "Byf", "Bzf"):
            A[k] = A[k] + rng.normal(0, 0.1, A[k].size)
    dev = torch.device("cuda:0")
    Ts = [{k: torch.from_numpy(A[k].copy()).to(dev) for k in ("U", "ent", "Bxf", "Byf", "Bzf")} for A in As]
    orc.p1_shell_exchange_oracle(orc.Grid(cfg), As)
    solver = cs.CubeSphere(cfg)
    try:
        solver.halo_exchange_p1(Ts)
        torch.cuda.synchronize()
    finally:
        solver.close()
    for s in range(6):
        for k in ("U", "ent", "Bxf", "Byf", "Bzf"):
            np.testing.assert_array_equal(Ts[s][k].cpu().numpy(), As[s][k], err_msg=f"patch {s} {k}")


def test_p1_sync_shared_gpu_matches_oracle():
    """cs_sync_shared_p1 against p1_shell_sync_oracle on six patches after
    their fluxes and edge fields: the fluctuations, entropy fluxes and edge
    fields exactly (means of the same two or three values), dU_dt to 1e-14 of
    each component's largest magnitude (the corrections of a corner zone's
    two shared faces are added in either order)."""
    N, NR, g = 8, 4, 4
    cfg, probs, As, grid = _shell_after_fluxes(N, NR, g, seed=5)
    faces, edges = cs.p1_shell_pairings(As, N, NR, g)
    dev = torch.device("cuda:0")
    keys = [k for k, _ in cs.P1PatchFluxes._fields_]
    Ts = [{k: torch.from_numpy(A[k].copy()).to(dev) for k in keys} for A in As]
    fd = torch.from_numpy(faces.view(np.uint8).copy()).to(dev)
    ed = torch.from_numpy(edges.view(np.uint8).copy()).to(dev)
    orc.p1_shell_sync_oracle(As, faces, edges)
    cs.sync_shared_p1(N, NR, g, Ts, fd, faces.size, ed, edges.size)
    torch.cuda.synchronize()
    for s in range(6):
        for k in keys:
            if k in ("xf_area", "yf_area", "zf_area", "v_volume"):
                continue
            G = Ts[s][k].cpu().numpy()
            if k == "dU_dt":
                bad = _compare({k: G}, {k: As[s][k]}, rtol=1e-14, rows={k: 8})
                assert not bad, (s, bad)
            else:
                np.testing.assert_array_equal(G, As[s][k], err_msg=f"patch {s} {k}")
