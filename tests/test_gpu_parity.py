"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Tolerances: ghost exchange is bit-exact (copies, sign flips and the fixed
(a + (-b)) * 0.5 averaging); floating-point results differ from the oracle
only by FMA contraction and libm ulps, so face fluxes and EMFs must agree to
1e-11 relative and a few SSP-RK3 steps to 1e-10 relative L1.
"""
import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _cases():
    return {
        # configuration C1's scheme on its own grid: 6 x 32^2 surface, WENO5 + HLL + SSP-RK3, Gaussian bump
        # with solid-body rotation and seeded density noise
        "C1_surf32_weno5_hll_rk3": cs.workload("C1")[:2],
        "surf16_hlld_blast": (cs.make_config(16, 1, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3), 2),
        "surf12_hll_rk2_g4": (cs.make_config(12, 1, 4, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2), 1),
        "shell10x4_hlld_grav": (cs.make_config(10, 4, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3,
                                               cs.RBC_PAD, cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0), 3),
        "shell8x3_hll_emfhll_log": (cs.make_config(8, 3, 3, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_HLL, cs.FP64, 3,
                                                   cs.RBC_PAD, cs.RMESH_LOG, 1.0, 2.0), 3),
        "surf8_pcm_wall": (cs.make_config(8, 1, 3, cs.RECON_PCM, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 3), 1),
        "shell8x3_imex_grav": (cs.make_config(8, 3, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2,
                                              cs.RBC_PAD, cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1), 3),
        "shell8x3_rot_grav": (cs.make_config(8, 3, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3,
                                             cs.RBC_PAD, cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0,
                                             omega=(0.1, -0.2, 0.4)), 3),
        "surf12_rot": (cs.make_config(12, 1, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3,
                                      omega=(0.0, 0.0, 0.5)), 1),
    }


CASES = _cases()


def _rel(a, b):
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d


@pytest.mark.parametrize("name", list(CASES))
def test_halo_bit_exact(name):
    cfg, prob = CASES[name]
    print(f"case {name}", flush=True)
    for p in (0, prob):
        st = cs.init_problem(cfg, p, seed=11)
        dev = cs.CubeSphere(cfg)
        dev.upload(st)
        dev.halo_exchange()
        got = dev.download()
        dev.close()
        ref = st.copy()
        orc.Grid(cfg).halo(ref)
        for nm, a, b in zip(["U", "ent", "Bx", "By", "Bz"], got.arrays(), ref.arrays()):
            bad = np.flatnonzero(a != b)
            assert bad.size == 0, f"{name} problem {p}: {nm} differs at {bad.size} entries, first {bad[:5]}"


@pytest.mark.parametrize("name", list(CASES))
def test_face_fluxes_and_emf(name):
    cfg, prob = CASES[name]
    print(f"case {name}", flush=True)
    st = cs.init_problem(cfg, prob, seed=3)
    G = orc.Grid(cfg)
    ref = st.copy()
    G.halo(ref)
    F_o, S_o = G.fluxes(ref)
    E_o = G.emf(ref, S_o)
    dev = cs.CubeSphere(cfg)
    dev.upload(st)
    dev.halo_exchange()
    dev.make_facial_fluxes()
    dev.make_edge_efields()
    for d in range(3):
        F, S = dev.fluxes(d)
        assert _rel(F, F_o[d]) < 1e-11, f"flux dir {d}: {_rel(F, F_o[d]):.3e}"
        assert _rel(S, S_o[d]) < 1e-11, f"speeds dir {d}: {_rel(S, S_o[d]):.3e}"
    E = dev.emf()
    for d in range(3):
        assert _rel(E[d], E_o[d]) < 1e-10, f"emf dir {d}: {_rel(E[d], E_o[d]):.3e}"
    dev.close()


@pytest.mark.parametrize("name", list(CASES))
def test_steps_match_oracle(name):
    cfg, prob = CASES[name]
    print(f"case {name}", flush=True)
    st = cs.init_problem(cfg, prob, seed=5)
    G = orc.Grid(cfg)
    dt0 = G.cfl(st, 0.0)
    ref = st.copy()
    dt_o = G.advance(ref, 3, dt0)
    dev = cs.CubeSphere(cfg)
    dev.upload(st)
    dev.dt = dt0
    dev.advance(3, use_graph=False)
    got = dev.download()
    dt_g = dev.dt
    dev.close()
    Ug, eg = got.interior()
    Uo, eo = ref.interior()
    rel = np.abs(Ug - Uo).sum() / np.abs(Uo).sum()
    assert rel < 1e-10, f"U rel L1 {rel:.3e}"
    assert rel_l1(eg, eo) < 1e-10
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    # owned faces (ghost layers are stage scratch; the surface-mode stage does
    # not refresh the radial ghosts it never reads)
    g, N, NR = cfg.nghost, cfg.ntp, cfg.nr
    owned = [(slice(g + 1, g + NR + 1), slice(g + 1, g + N + 1), slice(g, g + N + 1)),
             (slice(g + 1, g + NR + 1), slice(g, g + N + 1), slice(g + 1, g + N + 1)),
             (slice(g, g + NR + 1), slice(g + 1, g + N + 1), slice(g + 1, g + N + 1))]
    for a, b, sl in zip(got.views()[2:], ref.views()[2:], owned):
        assert _rel(a[(slice(None),) + sl], b[(slice(None),) + sl]) < 1e-9


def rel_l1(a, b):
    return np.abs(a - b).sum() / np.abs(b).sum()


@pytest.mark.parametrize("name", [n for n in CASES if n.startswith("shell")])
def test_pushed_ghosts_equal_exchange(name):
    """In a shell the zone and face updates write the lateral ghosts they
    feed; after whole steps (graph and eager) every lateral ghost zone and face
    must equal what a full exchange then copies into it, bit for bit (radial
    ghost layers, which the next stage's boundary kernel fills, are left out).
    Surface mode keeps the exchange launch at the start of each stage."""
    cfg, prob = CASES[name]
    print(f"case {name}", flush=True)
    g, NR = cfg.nghost, cfg.nr
    st = cs.init_problem(cfg, prob, seed=13)
    dev = cs.CubeSphere(cfg)
    dev.upload(st)
    dev.dt = 0.0
    dev.evaluate_timestep()
    dev.advance(2, use_graph=True)
    dev.advance(1, use_graph=False)
    pushed = dev.download()
    dev.halo_exchange()
    pulled = dev.download()
    dev.close()
    zk = slice(g, g + NR)              # zones k in [1, NR]
    fk = slice(g + 1, g + NR + 1)      # x- and y-faces k in [1, NR]
    zk_z = slice(g, g + NR + 1)        # z-faces k in [0, NR]
    a, b = pushed.views(), pulled.views()
    for nm, x, y in (("U", a[0][:, :, zk], b[0][:, :, zk]), ("ent", a[1][:, zk], b[1][:, zk]),
                     ("Bx", a[2][:, fk], b[2][:, fk]), ("By", a[3][:, fk], b[3][:, fk]),
                     ("Bz", a[4][:, zk_z], b[4][:, zk_z])):
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{name}: {nm} differs at {bad.size} entries, first {bad[:5]}"


def test_graph_replay_equals_eager():
    cfg, prob = CASES["surf16_hlld_blast"]
    st = cs.init_problem(cfg, prob, seed=9)
    outs = []
    for use_graph in (False, True):
        dev = cs.CubeSphere(cfg)
        dev.upload(st)
        dev.dt = 0.0
        dev.evaluate_timestep()
        for _ in range(3):           # several replays with eager work between them
            dev.advance(2, use_graph=use_graph)
            _ = dev.conservation_sums()
        outs.append(dev.download())
        dev.close()
    for a, b in zip(outs[0].arrays(), outs[1].arrays()):
        assert np.array_equal(a, b)


def test_divb_and_conservation():
    cfg, prob = cs.make_config(24, 1, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3), 2
    st = cs.init_problem(cfg, prob, seed=1)
    dev = cs.CubeSphere(cfg)
    dev.upload(st)
    s0 = dev.conservation_sums()
    d0 = dev.max_divb()
    dev.dt = 0.0
    dev.evaluate_timestep()
    dev.advance(20)
    s1 = dev.conservation_sums()
    d1 = dev.max_divb()
    dev.close()
    assert d0 < 1e-12 and d1 < 1e-12, (d0, d1)
    # mass and energy are conserved on the closed shell (walls) to roundoff
    assert abs(s1[0] - s0[0]) <= 1e-12 * abs(s0[0])
    assert abs(s1[4] - s0[4]) <= 1e-12 * abs(s0[4])


def test_mixed_precision_close_to_fp64():
    """A regression bound on the stage path's mixed form (FP32 WENO5 and
    Riemann solves, FP64 metric, CT, update and sums), not the benchmark's
    1e-6 gate: this form does not meet that gate (FP32 rounding of the
    primitives puts ~6e-8 p of noise into the momentum flux, large against
    the small momentum of the quiescent region), and the benchmark reports
    its gate as failed.  20 steps of the magnetised blast stay within 1e-5
    relative L1 of FP64."""
    base = dict(ntp=24, nr=1, nghost=3, recon=cs.RECON_WENO5, riemann=cs.RS_HLLD, emf=cs.EMF_LLF, rk_order=3)
    res = {}
    for prec in (cs.FP64, cs.MIXED):
        cfg = cs.make_config(precision=prec, **base)
        st = cs.init_problem(cfg, 1, seed=2)
        dev = cs.CubeSphere(cfg)
        dev.upload(st)
        dev.dt = 0.0
        dev.evaluate_timestep()
        dev.advance(20)
        res[prec] = dev.download()
        dev.close()
    U64, _ = res[cs.FP64].interior()
    U32, _ = res[cs.MIXED].interior()
    rel = np.abs(U32 - U64).sum() / np.abs(U64).sum()
    print(f"mixed vs fp64 rel L1 after 20 steps: {rel:.3e}")
    assert rel < 1e-5


def test_mixed_shell_imex_close_to_fp64():
    """The 3-D shell configuration's precision comparison at a small size
    (6x10^2x4, WENO5 + HLLD + CT, IMEX with implicit gravity, hydrostatic
    state with noise): after 10 steps density, energy and zone B of the mixed
    form agree with FP64 to 1e-6 relative L1 (measured 2e-8 to 1e-7) and div B
    stays at rounding in both; the momentum, nearly zero in hydrostatic
    balance, carries FP32 rounding of the pressure-gravity balance against its
    tiny magnitude (measured 6e-7) and is bounded at 1e-5.  Over the full
    configuration's grid and step count the momentum error grows past the
    1e-6 gate, which the benchmark reports."""
    base = dict(ntp=10, nr=4, nghost=3, recon=cs.RECON_WENO5, riemann=cs.RS_HLLD, emf=cs.EMF_LLF, rk_order=2,
                radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5, grav=1.0, imex=1)
    res, divb = {}, {}
    for prec in (cs.FP64, cs.MIXED):
        cfg = cs.make_config(precision=prec, **base)
        st = cs.init_problem(cfg, 3, seed=7)
        dev = cs.CubeSphere(cfg)
        dev.upload(st)
        dev.dt = 0.0
        dev.evaluate_timestep()
        dev.advance(10)
        res[prec] = dev.download()
        divb[prec] = dev.max_divb()
        dev.close()
    U64, _ = res[cs.FP64].interior()
    U32, _ = res[cs.MIXED].interior()
    rel = [float(np.abs(U32[:, c] - U64[:, c]).sum() / np.abs(U64[:, c]).sum()) for c in range(8)]
    print("mixed vs fp64 rel L1 per component after 10 steps:", " ".join(f"{r:.2e}" for r in rel))
    for c in (0, 4, 5, 6, 7):
        assert rel[c] < 1e-6, (c, rel)
    assert max(rel[1:4]) < 1e-5, rel
    assert divb[cs.FP64] < 1e-12 and divb[cs.MIXED] < 1e-12, divb


@pytest.mark.parametrize("name", ["surf16_hlld_blast", "shell10x4_hlld_grav"])
def test_autotune_keeps_results(name):
    """Workgroup sizes change only the schedule: the same steps give the same bits."""
    cfg, prob = CASES[name]
    print(f"case {name}", flush=True)
    st = cs.init_problem(cfg, prob, seed=21)
    outs = []
    for tune in (False, True):
        dev = cs.CubeSphere(cfg)
        dev.upload(st)
        dev.dt = 0.0
        dev.evaluate_timestep()
        if tune:
            sizes = dev.autotune(2)
            assert sizes["flux_xy"] in (64, 128, 256)
            dev.upload(st)
            dev.dt = 0.0
            dev.evaluate_timestep()
        dev.advance(3, use_graph=True)
        outs.append(dev.download())
        dev.close()
    for a, b in zip(outs[0].arrays(), outs[1].arrays()):
        assert np.array_equal(a, b)
