"""The reference's language on the drop-in boundary: a Fortran program using
the ISO_C_BINDING module (cubesphere_amd/fortran/) runs one SSP-RK3 step in
the driver's array layout through the C-ABI; its zone value must equal the
value the Python host path produces for the same inputs and call sequence."""
import os
import re
import subprocess

import numpy as np
import pytest

import cubesphere_amd as cs

pytestmark = pytest.mark.gpu
DEMO = os.path.join(os.path.dirname(cs.__file__), "cs_dropin_demo")


def test_fortran_dropin_matches_python_path():
    if not os.path.exists(DEMO):
        pytest.skip("Fortran demo not built (no flang)")
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"rho\(1,1,1,panel 3\) =\s*(\S+)\s+mass drift =\s*(\S+)\s+divB =\s*(\S+)", r.stdout)
    assert m, r.stdout
    rho_f, drift, divb = (float(x) for x in m.groups())
    cfg = cs.make_config(16, 1, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 3, cs.RBC_WALL, 0,
                         1.0, 1.0 + 2.0 * np.arctan(1.0) / 16)
    st = cs.init_problem(cfg, 2, seed=7)
    dev = cs.CubeSphere(cfg)
    dev.upload(st)
    dev.dt = 0.0
    dev.evaluate_timestep()
    for irk in (1, 2, 3):
        dev.rk_stage(irk)
    got = dev.download()
    dev.close()
    U = got.views()[0]
    g = cfg.nghost
    rho_p = U[2, 0, g, g, g]          # panel 3, zone (1,1,1)
    assert abs(rho_f - rho_p) <= 1e-4 * abs(rho_p)   # printed with 5 significant digits
    assert abs(drift) < 1e-12 and divb < 1e-12
