"""smdlt_mc_sweep_noalloc (the allocation-free chargeHop sweep) against
smdlt_mc_sweep on identical state: same consumed counts, logs, charge indices
and displacement sums, bit for bit, on two calls in a row, the second with a
smaller input; and its size limit is enforced on the host."""
import ctypes as C

import numpy as np
import pytest
from helpers import S, random_walk

pytestmark = pytest.mark.gpu


def _pair(mode, nb=30, nc=4, nq=50, L=22.0, flow=S.FLOW_SHEAR):
    """Two contexts holding the same chains and charges."""
    pos = random_walk(nb, nc, L, 21)
    track = np.random.default_rng(5).choice(nb * nc, nq, replace=False).astype(np.int32)
    sign = np.where(np.arange(nq) % 4 == 0, -1, 1).astype(np.int32)
    out = []
    for _ in range(2):
        p = S.default_params(nb=nb, nc=nc, L=L, flow_type=flow, edot=0.3, dt=1e-3, lambda_d=5.0,
                             lambda_b=3.59, barrier=1.0, barrier2=1.5)
        c = S.Context(p)
        c.upload_positions(*[np.ascontiguousarray(pos[:, k]) for k in range(3)])
        c.remap(40)
        c.upload_charges([nq], track, sign)
        out.append(c)
    return out


def _same(a, b, mode, u):
    ca, la = a.mc_sweep(mode, u)
    cb, lb = b.mc_sweep(mode, u, noalloc=True)
    assert np.array_equal(ca, cb)
    assert np.array_equal(la, lb)
    ta, oa = a.download_charges()
    tb, ob = b.download_charges()
    assert np.array_equal(ta, tb) and np.array_equal(oa, ob)
    assert np.array_equal(a.charge_displacements(), b.charge_displacements())
    return ca


@pytest.mark.parametrize("mode", [S.MC_REFERENCE, S.MC_MH_UNIFORM])
def test_noalloc_sweep_matches_allocating_sweep(mode):
    a, b = _pair(mode)
    rs = S.RefStream(830062920)
    if mode != S.MC_REFERENCE:
        # the MH lists' capacity comes from an allocating sweep: without one the
        # allocation-free call refuses; then one identical sweep on both
        with pytest.raises(S.SmdltError):
            b.mc_sweep(mode, rs.uniforms(3 * 50)[None, :], noalloc=True)
        u0 = rs.uniforms(3 * 50)[None, :]
        a.mc_sweep(mode, u0)
        b.mc_sweep(mode, u0)
    u1 = rs.uniforms(4 * 50)[None, :]
    cons = _same(a, b, mode, u1)
    assert cons[0] > 0
    # second call, smaller input: the array still holds the first call's
    # uniforms past the new length; only the new ones may be read
    u2 = rs.uniforms(3 * 50)[None, :]
    _same(a, b, mode, u2)
    # device generator (no host uniforms) on both
    _same(a, b, mode, None)
    a.close()
    b.close()


def test_noalloc_rejects_oversize_input():
    a, b = _pair(S.MC_REFERENCE, nq=10)
    cap = S.lib().smdlt_mc_sweep_noalloc
    nu = (8 << 20) // 4  # the uniforms alone fill the array; with the counts they exceed it
    u = np.full(nu, 0.5, np.float32)
    cons = np.zeros(1, np.int32)
    log = np.zeros(60, np.int32)
    code = cap(b.h, S.MC_REFERENCE, u.ctypes.data_as(C.POINTER(C.c_float)), nu,
               cons.ctypes.data_as(C.POINTER(C.c_int)), log.ctypes.data_as(C.POINTER(C.c_int)))
    assert code == -1  # SMDLT_E_ARG, before any device work
    # the context is still usable afterwards, and matches the allocating sweep
    rs = S.RefStream(981272199)
    _same(a, b, S.MC_REFERENCE, rs.uniforms(40)[None, :])
    a.close()
    b.close()
