"""The P1 geometry built on the GPU (cs_build_geometry, through the C-ABI)
against the C oracle, which tests/test_cpu_geometry.py pins against the
reference's own routines.  The kernels round like the oracle (no contraction,
IEEE division and square root), so vertices, masks, scale lengths and the face
quadrature agree to the last bit or an ulp (1e-15 relative), the volume
quadrature to 1e-13 relative, and the rounding-level first moments to 1e-13
absolute."""
import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CASES = {
    "shell6x3_g3": cs.make_config(6, 3, 3, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.5),
    "shell12x5_g4": cs.make_config(12, 5, 4, radial_bc=cs.RBC_PAD, rmin=1.0, rmax=1.3),
    "shell7x3_g3_log": cs.make_config(7, 3, 3, radial_bc=cs.RBC_PAD, radial_mesh=cs.RMESH_LOG, rmin=1.0, rmax=2.0),
    "surface16": cs.make_config(16, 1, 3),
}


def _rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a - b).max())


@pytest.mark.parametrize("name", list(CASES))
def test_geometry_matches_oracle(name):
    cfg = CASES[name]
    print(f"case {name}", flush=True)
    got, ms = cs.build_geometry(cfg)
    ref = orc.quad_geometry(cfg)
    assert np.array_equal(got["mask"], ref["mask"])
    for k in ("vert", "scale", "v_pts"):
        assert _rel(got[k], ref[k]) <= 1e-15, k
    for f in ("xf", "yf", "zf"):
        for k in cs.GEOMETRY_FACE_KEYS:
            assert _rel(got[f][k], ref[f][k]) <= 1e-15, f"{f}.{k}"
    for f in ("xb", "yb", "zb"):
        for k in cs.GEOMETRY_FACE_KEYS[:3]:
            assert _rel(got[f][k], ref[f][k]) <= 1e-15, f"{f}.{k}"
    for k in ("v_wts", "v_tan", "v_vol", "v_cen"):
        assert _rel(got[k], ref[k]) < 1e-13, k
    assert np.abs(got["moments"] - ref["moments"]).max() < 1e-13
    assert all(t >= 0.0 for t in ms)
