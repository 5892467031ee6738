"""cs_limit_all_slopes_p1_noalloc (LIMIT_ALL_SLOPES_P1 with its work memory in
the library's fixed per-GPU array, nothing allocated) against
cs_limit_all_slopes_p1 through the C-ABI: two calls in a row, the second on a
smaller patch, each bit-identical with the allocating entry point on the same
inputs (the same kernels in the same order); a patch larger than the array is
refused before anything runs."""
import ctypes

import numpy as np
import pytest

import cubesphere_amd as cs
from oracle import oracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _run(prob, A, noalloc):
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(A[k]).to(dev) for k in cs.P1_ARRAYS}
    vmin, vmax = cs.limit_all_slopes_p1(prob["N"], prob["NR"], prob["g"], prob, T, noalloc=noalloc)
    torch.cuda.synchronize()
    out = {k: T[k].cpu().numpy() for k in cs.P1_ARRAYS}
    out["V_min_local"], out["V_max_local"] = vmin, vmax
    return out


def test_noalloc_equals_allocating_entry_point():
    # the first patch has noise enough that zones lose their slopes and the
    # positivity replay runs; the second is smaller and smooth
    cases = [(12, 5, 1, 2, True, 1.0), (8, 3, 0, 1, False, 0.0)]
    for N, NR, sector, seed, shock, noise in cases:
        print(f"\nnoalloc case N={N} NR={NR} sector={sector} seed={seed} noise={noise}", flush=True)
        prob, A0 = orc.p1_problem(N, NR, 4, sector=sector, seed=seed, shock=shock, noise=noise)
        ref = _run(prob, orc.p1_copy(A0), noalloc=False)
        got = _run(prob, orc.p1_copy(A0), noalloc=True)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), k
    # the oracle agrees too (the allocating entry point's own tolerance)
    O = orc.p1_oracle(prob, orc.p1_copy(A0))
    for k in ("V", "U", "ent", "Bxf_q"):
        scale = max(np.abs(O[k]).max(), 1e-300)
        assert np.abs(got[k] - O[k]).max() <= 1e-10 * scale, k


def test_noalloc_refuses_a_patch_beyond_its_array():
    # a patch whose work memory exceeds the fixed array: refused on the host
    # before any launch (the arrays are never touched)
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    s = cs.P1Patch()
    s.ntp, s.nr, s.nghost = 400, 200, 4
    for q in range(4):
        s.itri[q] = 1
    for k in cs.P1_SCALARS:
        setattr(s, k, 1.0)
    s.gamma, s.smallnum, s.rhofloor, s.prsfloor = 5.0 / 3.0, 1e-12, 1e-12, 1e-9
    for k in cs.P1_ARRAYS:
        setattr(s, k, dummy.data_ptr())
    rc = cs.lib().cs_limit_all_slopes_p1_noalloc(ctypes.byref(s), None)
    assert rc == -1
    assert "work memory" in cs.last_error()
