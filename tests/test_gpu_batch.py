"""Batched shells through the C-ABI: each phase of a stage is one launch over
all shells.  The batched kernels run the same device functions on the same
data as the single-shell path, so after a few steps every shell must be bit
for bit the state the single-shell path produces (whose parity with the
oracle is tested in test_gpu_parity.py), with explicit and implicit gravity,
in FP64 and in the mixed form (the 48-shell configuration's two legs), and
after the autotuner has changed the workgroup sizes."""
import numpy as np
import pytest

import cubesphere_amd as cs

pytestmark = pytest.mark.gpu


def _shells(imex, precision=cs.FP64):
    kw = dict(recon=cs.RECON_WENO5, riemann=cs.RS_HLLD, emf=cs.EMF_LLF, precision=precision,
              rk_order=2 if imex else 3, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5, grav=1.0, imex=imex)
    cfgs = [cs.make_config(8, 3, **kw), cs.make_config(10, 4, **kw), cs.make_config(12, 2, **kw),
            cs.make_config(9, 5, riemann=cs.RS_HLL, **{k: v for k, v in kw.items() if k != "riemann"})]
    states = [cs.init_problem(c, 2 + (i % 2), seed=40 + i) for i, c in enumerate(cfgs)]
    return cfgs, states


def _one_by_one(cfgs, states, nsteps):
    out = []
    for c, st in zip(cfgs, states):
        dev = cs.CubeSphere(c)
        dev.upload(st)
        dev.dt = 0.0
        dev.evaluate_timestep()
        dev.advance(nsteps, use_graph=False)
        out.append((dev.download(), dev.dt))
        dev.close()
    return out


@pytest.mark.parametrize("imex,tune,precision", [(0, False, cs.FP64), (1, False, cs.FP64), (0, True, cs.FP64),
                                                (1, True, cs.FP64), (0, False, cs.MIXED), (1, True, cs.MIXED)])
def test_batch_equals_single_shells(imex, tune, precision):
    print(f"batch imex={imex} autotune={tune} precision={precision}", flush=True)
    cfgs, states = _shells(imex, precision)
    ref = _one_by_one(cfgs, states, 3)
    B = cs.Batch(cfgs)
    for sh, st in zip(B.shells, states):
        sh.upload(st)
        sh.dt = 0.0
    B.evaluate_timestep()
    if tune:
        sizes = B.autotune(2)
        assert set(sizes.values()) <= {64, 128, 256}
        for sh, st in zip(B.shells, states):   # the tuner ran stages on the scratch buffers only
            sh.upload(st)
            sh.dt = 0.0
        B.evaluate_timestep()
    B.advance(1, use_graph=False)
    B.advance(2, use_graph=True)
    for n, (sh, (st_ref, dt_ref)) in enumerate(zip(B.shells, ref)):
        got = sh.download()
        for name, a, b in zip(["U", "ent", "Bx", "By", "Bz"], got.arrays(), st_ref.arrays()):
            assert np.array_equal(a, b), f"shell {n}: {name} differs"
        assert sh.dt == dt_ref
    B.close()


def test_batch_with_surface_shell():
    """A batch holding a surface-mode shell (one layer between walls) beside
    3-D shells: the surface shell's stages need the lateral exchange, which
    the batch then runs as its own phase.  The batch advances the surface
    shell with the general zone update (radial fluxes through the walls from
    the mirrored ghosts) where the single-shell path fuses the wall fluxes, so
    it agrees to rounding (1e-10 relative L1), not bit for bit; the 3-D
    shells stay bit-identical."""
    print("batch with a surface shell", flush=True)
    kw = dict(recon=cs.RECON_WENO5, riemann=cs.RS_HLLD, emf=cs.EMF_LLF, precision=cs.FP64, rk_order=3)
    cfgs = [cs.make_config(12, 1, **kw),
            cs.make_config(8, 3, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5, grav=1.0, **kw)]
    states = [cs.init_problem(cfgs[0], 2, seed=50), cs.init_problem(cfgs[1], 3, seed=51)]
    ref = _one_by_one(cfgs, states, 3)
    B = cs.Batch(cfgs)
    for sh, st in zip(B.shells, states):
        sh.upload(st)
        sh.dt = 0.0
    B.evaluate_timestep()
    B.advance(1, use_graph=False)
    B.advance(2, use_graph=True)
    got = [sh.download() for sh in B.shells]
    B.close()
    U, _ = got[0].interior()
    Ur, _ = ref[0][0].interior()
    rel = float(np.abs(U - Ur).sum() / np.abs(Ur).sum())
    assert rel < 1e-10, f"surface shell: rel L1 {rel:.3e}"
    for name, a, b in zip(["U", "ent", "Bx", "By", "Bz"], got[1].arrays(), ref[1][0].arrays()):
        assert np.array_equal(a, b), f"3-D shell: {name} differs"
