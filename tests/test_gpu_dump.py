"""Checkpoint / restart of a handle (cs_write_dump, cs_read_dump): the state
goes between device memory and the file through pinned staging; the file
holds U^n bit for bit, and a run restarted from it continues as the run that
wrote it.  The owned zones and faces are compared (ghost layers are stage
scratch, refreshed by the restart's upload).

They continue bit for bit where nothing but U^n carries over between steps.
Two schemes carry more, so their restart agrees to rounding only (1e-14 of
each variable's largest magnitude; a vector component, of its vector's):
  * the WENO5 surface stage keeps the primitives its zone update produced,
    and CONS_TO_PRIM under its floors is not idempotent, so the primitives a
    restart recomputes from U can differ in the last bit;
  * the P1 scheme with IMEX reads U^n's face B in its second stage without
    the exchange having averaged the two copies of each panel-edge face (the
    first stage reads the implicit stage's copy), while a restart's upload
    averages them."""
import numpy as np
import pytest

import cubesphere_amd as cs

pytestmark = pytest.mark.gpu


def _cases():
    # name: (config, problem, bit for bit)
    return {
        "shell8x3_weno5_hlld_imex": (cs.make_config(8, 3, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2,
                                                    cs.RBC_PAD, cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1), 3,
                                     True),
        "surf12_weno5_hll_rk3": (cs.make_config(12, 1, 3, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 3), 1,
                                 False),
        "shell8x3_p1adp_hlld_rk2_grav": (cs.make_config(8, 3, 4, cs.RECON_P1ADP, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2,
                                                        cs.RBC_PAD, cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0), 3, True),
        "C5_scheme_shell8x3_p1adp": (cs.make_config(8, 3, 4, cs.RECON_P1ADP, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2,
                                                    cs.RBC_PAD, cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1), 3,
                                     False),
        "C2_scheme_surf12_p1adp": (cs.make_config(12, 1, 4, cs.RECON_P1ADP, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3), 2,
                                   True),
    }


def _owned_faces(cfg, q):
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    nf, nfr = N + 2 * g + 1, NR + 2 * g + 1
    m = np.zeros((nfr, nf, nf), bool)
    i0, j0, k0 = (0 if q == 0 else 1), (0 if q == 1 else 1), (0 if q == 2 else 1)
    m[k0 + g:NR + 1 + g, j0 + g:N + 1 + g, i0 + g:N + 1 + g] = True
    return np.tile(m.ravel(), 6)


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_restart_continues_the_run(name, tmp_path):
    cfg, prob, exact = CASES[name]
    print(f"case {name}", flush=True)
    st = cs.init_problem(cfg, prob, seed=17)
    p = tmp_path / "d.csd"
    a = cs.CubeSphere(cfg)
    try:
        a.upload(st)
        a.dt = 0.0
        a.evaluate_timestep()
        a.advance(2)
        a.write_dump(p, time=0.125, filenum=4, step=2)
        dt_at_dump = a.dt
        mid = a.download()
        a.advance(2)
        ref = a.download()
        dt_ref = a.dt
    finally:
        a.close()
    # the file holds exactly U^n of the handle
    host, m = cs.read_dump(p, cfg)
    for x, y in zip(host.arrays(), mid.arrays()):
        assert np.array_equal(x, y)
    assert (m.time, m.filenum, m.step, m.dtcur) == (0.125, 4, 2, dt_at_dump)
    b = cs.CubeSphere(cfg)
    try:
        m2 = b.read_dump(p)
        assert m2.step == 2 and b.dt == dt_at_dump
        b.advance(2)
        got = b.download()
        dt_got = b.dt
    finally:
        b.close()
    Ug, eg = got.interior()
    Ur, er = ref.interior()
    # vector components scaled by their vector's largest component
    group = {1: slice(1, 4), 2: slice(1, 4), 3: slice(1, 4), 5: slice(5, 8), 6: slice(5, 8), 7: slice(5, 8)}
    pairs = [(f"U[{c}]", Ug[:, c], Ur[:, c], np.abs(Ur[:, group.get(c, slice(c, c + 1))]).max()) for c in range(8)]
    pairs += [("ent", eg, er, np.abs(er).max())]
    bmax = max(np.abs(b[_owned_faces(cfg, q)]).max() for q, b in enumerate(ref.arrays()[2:]))
    pairs += [(nm, x[_owned_faces(cfg, q)], y[_owned_faces(cfg, q)], bmax)
              for q, (nm, x, y) in enumerate(zip(("Bx", "By", "Bz"), got.arrays()[2:], ref.arrays()[2:]))]
    for nm, x, y, scale in pairs:
        if exact:
            bad = np.flatnonzero(x != y)
            assert bad.size == 0, f"{nm} differs at {bad.size} entries, first {bad[:5]}"
        else:
            assert np.abs(x - y).max() <= 1e-14 * max(scale, 1e-300), (nm, np.abs(x - y).max(), scale)
    if exact:
        assert dt_got == dt_ref
    else:
        assert abs(dt_got - dt_ref) <= 1e-14 * dt_ref


def test_refused_file_leaves_the_handle(tmp_path):
    cfg, prob, _ = CASES["shell8x3_weno5_hlld_imex"]
    other = cs.make_config(10, 3, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                           cs.RMESH_UNIFORM, 1.0, 1.5)
    p = tmp_path / "o.csd"
    cs.write_dump(p, cs.init_problem(other, 1, seed=2))
    h = cs.CubeSphere(cfg)
    try:
        h.upload(cs.init_problem(cfg, prob, seed=3))
        h.dt = 1e-3
        h.advance(1)
        before = h.download()
        with pytest.raises(cs.CsError) as e:
            h.read_dump(p)
        assert e.value.code == cs.E_INVALID
        after = h.download()
    finally:
        h.close()
    for x, y in zip(before.arrays(), after.arrays()):
        assert np.array_equal(x, y)
