"""GPU checks of the five benchmark configurations against the CPU oracle:

- the small FENE chain (1 x 100 beads, L = 60, FP64): M.F, M^1/2 z and one
  step against the dense Ewald matrix and its symmetric square root;
- the semidilute shear system (8000 beads) and the planar-extension system
  (50,000 beads): M.F against the oracle's matrix-free Ewald sum at full size,
  and M^1/2 (M^1/2 z) = M z through two Lanczos solves;
- the uneven replica ensemble: a seeded sample of its replicas (both flows) in
  one batched context, each replica's M.F against the oracle and M^1/2 z
  against the dense square root, and the full 512-replica ensemble through the
  graph path with property checks;
- the charge-hopping system (20 x 300 beads, 2400 charges, hop radius 3a): the
  reference chargeHop sequence and the Metropolis-Hastings sequence against the
  oracle's;
- overlapping beads (pairs down to r = 0.5) in mixed precision, and an MC
  acceptance exactly at the tie exp(0) = 1.

Measured errors are printed (pytest -s shows them)."""
import numpy as np
import pytest
from helpers import O, S, cell_from_lattice_row, random_walk

import smdlt_workloads as W

pytestmark = pytest.mark.gpu


def xyz(pos):
    return [np.ascontiguousarray(pos[:, k]) for k in range(3)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def wrap_frame(ctx, t):
    """Device positions and per-replica oracle cells in the frame M is taken in at step t."""
    ctx.remap(t)
    x, y, z = ctx.download_positions()
    rows = ctx.update_lattice(t)
    L = [ctx.params[r].L for r in range(ctx.nrep)]
    return np.stack([x, y, z], 1), [cell_from_lattice_row(rows[r], L[r]) for r in range(ctx.nrep)]


def dense_sqrt(M, z):
    w, Q = np.linalg.eigh(M)
    return Q @ (np.sqrt(np.clip(w, 0, None)) * (Q.T @ z))


# ------------------------------------------------------------------ 1 x 100 FENE chain
def c1_ctx(**kw):
    p = S.default_params(nb=100, nc=1, L=60.0, flow_type=S.FLOW_SHEAR, edot=10.0 / 175.0, dt=1e-3, spring_fene=1,
                         kappas=30.0, qmax=10.0, precision=S.PREC_FP64, lanczos_tol=1e-6, lanczos_maxit=30,
                         pbc_compat=1, **kw)
    ctx = S.Context(p)
    pos = W.chains(100, 1, 60.0, 8)
    ctx.upload_positions(*xyz(pos))
    return ctx


def test_small_fene_chain_mobility_and_sqrt_match_dense():
    ctx = c1_ctx(mobility_tol=1e-7)
    pos, cells = wrap_frame(ctx, 2000)
    M = O.ewald_dense(cells[0], pos, 0.4)
    rng = np.random.default_rng(1)
    v = rng.normal(size=pos.size)
    e_mf = rel(ctx.mobility_apply(v), M @ v)
    z = rng.normal(size=pos.size)
    y, it, res, code = ctx.brownian(z, 1e-6, 30)
    ys, st, lo, hi = O.sym_sqrt_apply(M, z)
    e_sq = rel(y, ys)
    print(f"1x100 FENE chain, FP64: M.F err {e_mf:.2e}, M^1/2 z err {e_sq:.2e} ({it[0]} Lanczos iterations)")
    assert code == 0 and e_mf < 1e-5 and e_sq < 1e-5
    ctx.close()


def test_small_fene_chain_step_matches_dense_update():
    # one step with host noise z: dx = dt M F + sqrt(2 dt) M^1/2 z + dt (edot y, 0, 0)
    ctx = c1_ctx(mobility_tol=1e-7)
    t, dt, edot = 2000, 1e-3, 10.0 / 175.0
    pos, cells = wrap_frame(ctx, t)
    lat = O.lattice(1, 60.0, edot, dt, 1, t)
    rx, ry, rz = xyz(pos)
    nl, lst = O.verlet(lat, rx, ry, rz, 6.0)
    st, fb, _ = O.bond_force(lat, 100, 1, rx, ry, rz, 1, 30.0, 10.0)
    fl, _ = O.lj_force(lat, rx, ry, rz, nl, lst, 25.0, 0.1, 0)
    F = (fb + fl).reshape(-1)
    M = O.ewald_dense(cells[0], pos, 0.4)
    z = np.random.default_rng(2).normal(size=pos.size)
    ys, _, _, _ = O.sym_sqrt_apply(M, z)
    d = (dt * (M @ F) + np.sqrt(2 * dt) * ys).reshape(-1, 3)
    d[:, 0] += dt * edot * ry
    ctx.upload_positions(rx, ry, rz)
    ctx.step(t, z)
    g = np.stack(ctx.download_positions(), 1)
    e = rel(g - pos, d)
    print(f"1x100 FENE chain, FP64: step displacement err {e:.2e}")
    assert e < 1e-5
    ctx.close()


# ------------------------------------------------------------------ full-size systems
@pytest.fixture(scope="module")
def semidilute():
    ctx = S.Context(W.c2_params(S.PREC_MIXED))
    pos = W.chains(200, 40, 80.0, 1)
    ctx.upload_positions(*xyz(pos))
    ctx.upload_charges([4000], *W.c2_charges(8000))
    ctx.run(0, 2, use_graph=True)
    yield ctx
    ctx.close()


def test_semidilute_mobility_matches_matrix_free_ewald(semidilute):
    ctx = semidilute
    pos, cells = wrap_frame(ctx, 2)
    v = np.random.default_rng(3).normal(size=pos.size)
    e = rel(ctx.mobility_apply(v), O.ewald_matvec(cells[0], pos, v))
    print(f"8000-bead shear: M.F err vs matrix-free Ewald oracle {e:.2e}")
    assert e < 1e-5


def test_semidilute_sqrt_squares_to_oracle_product(semidilute):
    # M^1/2 (M^1/2 z) = M z, with M z from the oracle
    ctx = semidilute
    pos, cells = wrap_frame(ctx, 2)
    z = np.random.default_rng(4).normal(size=pos.size)
    y, it, res, code = ctx.brownian(z, 1e-6, 128)
    y2, it2, res2, code2 = ctx.brownian(y, 1e-6, 128)
    e = rel(y2, O.ewald_matvec(cells[0], pos, z))
    print(f"8000-bead shear: M^1/2 M^1/2 z vs oracle M z err {e:.2e} ({it[0]}, {it2[0]} iterations)")
    assert code == 0 and code2 == 0 and e < 1e-5


def test_planar_extension_50k_mobility_matches_matrix_free_ewald():
    nb, nc, L = 250, 200, 160.0
    p = S.default_params(nb=nb, nc=nc, L=L, flow_type=S.FLOW_EXTENSION, edot=20.0 / 175.0, dt=2e-4,
                         kappas=200.0, qmax=2.0, epsilon=0.1, lj_ext_form=1, precision=S.PREC_MIXED,
                         lanczos_tol=1e-6, lanczos_maxit=128)
    ctx = S.Context(p)
    pos = W.to_lab_frame(ctx, [{"flow": S.FLOW_EXTENSION}], [W.chains(nb, nc, L, 4)])[0]
    ctx.upload_positions(*xyz(pos))
    pos, cells = wrap_frame(ctx, 1300)
    v = np.random.default_rng(5).normal(size=pos.size)
    e = rel(ctx.mobility_apply(v), O.ewald_matvec(cells[0], pos, v))
    print(f"50,000-bead planar extension: M.F err vs matrix-free Ewald oracle {e:.2e}")
    assert e < 1e-5
    ctx.close()


# ------------------------------------------------------------------ uneven replica ensemble
def c4_sample(k=16, nmax=3000, seed=3):
    specs = W.c4_specs(512)
    rng = np.random.default_rng(seed)
    idx = [r for r in rng.permutation(512) if specs[r]["nb"] * specs[r]["nc"] <= nmax]
    sh = [r for r in idx if specs[r]["flow"] == S.FLOW_SHEAR][:k // 2]
    ex = [r for r in idx if specs[r]["flow"] == S.FLOW_EXTENSION][:k // 2]
    return [specs[r] for r in sorted(sh + ex)]


def test_replica_sample_each_replica_matches_oracle():
    specs = c4_sample()
    ctx = S.Context(W.c4_params(specs, S.PREC_MIXED))
    pos = W.to_lab_frame(ctx, specs, W.grow_chains([(s["nb"], s["nc"], s["L"]) for s in specs], 6))
    allp = np.concatenate(pos)
    ctx.upload_positions(*xyz(allp))
    ctx.run(0, 2, use_graph=True)
    rng = np.random.default_rng(6)
    order = np.argsort([s["nb"] * s["nc"] for s in specs])
    errs, serrs = [], []
    # t = 2: the state the steps left; t = 7919: a different strain phase for
    # every replica (the lattice jumps there without the flow deforming the
    # chains, so images can overlap; M.F only)
    for t in (2, 7919):
        pw, cells = wrap_frame(ctx, t)
        v = rng.normal(size=pw.size)
        u = ctx.mobility_apply(v)
        if t == 2:
            z = rng.normal(size=pw.size)
            y, it, res, code = ctx.brownian(z, 1e-6, 128)
            assert code == 0
        off = 0
        for r, s in enumerate(specs):
            n = s["nb"] * s["nc"]
            sl = slice(3 * off, 3 * (off + n))
            p = pw[off:off + n]
            errs.append(rel(u[sl], O.ewald_matvec(cells[r], p, v[sl])))
            if t == 2 and r in order[:4]:  # dense square roots for the four smallest replicas
                Md = O.ewald_dense(cells[r], p, O.ewald_dense_alpha(cells[r]))
                serrs.append(rel(y[sl], dense_sqrt(Md, z[sl])))
                print(f"replica {r} ({n} beads): dense square root done", flush=True)
            off += n
    print("replica sample: sizes", [s["nb"] * s["nc"] for s in specs])
    print("replica sample: M.F err per replica", ["%.1e" % e for e in errs])
    print("replica sample: M^1/2 z err (4 smallest)", ["%.1e" % e for e in serrs])
    assert max(errs) < 1e-5 and max(serrs) < 1e-5
    ctx.close()


def test_full_ensemble_512_replicas_graph_steps():
    specs = W.c4_specs(512)
    pos = W.grow_chains_grouped([(s["nb"], s["nc"], s["L"]) for s in specs], 2)
    print("512 replicas: chains grown", flush=True)
    ctx = S.Context(W.c4_params(specs, S.PREC_MIXED))
    pos = W.to_lab_frame(ctx, specs, pos)
    allp = np.concatenate(pos)
    assert len(allp) > 2_500_000
    ctx.upload_positions(*xyz(allp))
    assert ctx.run(0, 3, use_graph=True) == 0
    print("512 replicas: 3 graph steps done", flush=True)
    it, status, change = ctx.lanczos_status()
    assert np.all(status == 1) and it.max() < 128
    x, y, z = ctx.download_positions()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(y)) and np.all(np.isfinite(z))
    # ||M^1/2 z||^2 = z.M.z on three replicas of different sizes and flows
    n = np.array([s["nb"] * s["nc"] for s in specs])
    off = np.concatenate([[0], np.cumsum(n)])
    ctx.remap(3)
    for r in (int(np.argmin(n)), int(np.argmax(n)), 257):
        ctx.select_replicas(r, r + 1)
        zz = np.zeros(3 * len(allp))
        zz[3 * off[r]:3 * off[r + 1]] = np.random.default_rng(r).normal(size=3 * n[r])
        yy, itr, res, code = ctx.brownian(zz, 1e-6, 128)
        mz = ctx.mobility_apply(zz)
        sl = slice(3 * off[r], 3 * off[r + 1])
        assert code == 0
        assert yy[sl] @ yy[sl] == pytest.approx(zz[sl] @ mz[sl], rel=1e-5)
    ctx.select_replicas(0, 512)
    ctx.close()


# ------------------------------------------------------------------ charge hopping, 20 x 300
@pytest.fixture(scope="module")
def hop_system():
    nb, nc, L, t = 300, 20, 60.0, 40
    pos = W.chains(nb, nc, L, 5)
    rng = np.random.default_rng(9)
    track = rng.choice(nb * nc, 2400, replace=False).astype(np.int32)
    return nb, nc, L, t, pos, track


def hop_ctx(hs, **kw):
    nb, nc, L, t, pos, track = hs
    p = S.default_params(nb=nb, nc=nc, L=L, flow_type=S.FLOW_SHEAR, edot=50.0 / 175.0, dt=5e-4, lambda_b=3.59,
                         lambda_d=5.0, hop_radius=3.0, pbc_compat=1, **kw)
    ctx = S.Context(p)
    ctx.upload_positions(*xyz(pos))
    ctx.remap(t)
    lat = O.lattice(1, L, 50.0 / 175.0, 5e-4, 1, t)
    return ctx, lat, xyz(np.stack(ctx.download_positions(), 1))


def test_charge_hopping_reference_sweep_matches_oracle(hop_system):
    nb, nc, L, t, pos, track = hop_system
    ctx, lat, (rx, ry, rz) = hop_ctx(hop_system, barrier=1.0, barrier2=2.0)
    ox, oy, oz = O.apply_pbc(lat, *xyz(pos))
    assert np.array_equal(ox, rx) and np.array_equal(oy, ry) and np.array_equal(oz, rz)
    ctx.upload_charges([len(track)], track)
    mc = O.ChargeHop(nb, nc, track)
    u = O.Ran1(830062920).uniforms(3 * len(track))
    cons, log = ctx.mc_sweep(S.MC_REFERENCE, u[None, :])
    st, cur, olog, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, 1.0, 2.0, u)
    tr, old = ctx.download_charges()
    print(f"2400-charge chargeHop sweep: {int(olog[:, 4].sum())} accepted, {cur} uniforms")
    assert st == 0 and cons[0] == cur
    assert np.array_equal(tr, mc.track)
    assert np.array_equal(log[:, 1], olog[:, 1]) and np.array_equal(log[:, 4:], olog[:, 4:])
    keep = olog[:, 3] >= 0
    assert np.array_equal(log[keep, 2:4], olog[keep, 2:4])
    ctx.close()


def test_charge_hopping_mh_sweeps_match_oracle(hop_system):
    nb, nc, L, t, pos, track = hop_system
    ctx, lat, (rx, ry, rz) = hop_ctx(hop_system, barrier=0.5, barrier2=1.0)
    nq = len(track)
    sign = np.where(np.arange(nq) < nq // 2, 1, -1).astype(np.int32)
    ctx.upload_charges([nq], track, sign)
    op = O.mh_params(nb * nc, nb, hop_radius=3.0, feature_radius=4.0, floor_eps=0.05, lambda_b=3.59,
                     lambda_d=5.0, barrier=0.5, barrier2=1.0)
    qbead = np.zeros(nb * nc, np.int32)
    qbead[track] = sign
    otrack = track.copy()
    rng = np.random.default_rng(10)
    for sweep in range(3):
        u = rng.random(3 * nq).astype(np.float32)
        cons, log = ctx.mc_sweep(S.MC_MH_UNIFORM, u[None, :])
        acc, olog, _ = O.mh_sweep(lat, op, rx, ry, rz, otrack, sign, qbead, None, u)
        assert cons[0] == 3 * nq and np.array_equal(log, olog)
        assert np.array_equal(ctx.download_charges()[0], otrack)
    print(f"2400-charge MH sweeps: last sweep accepted {acc}")
    ctx.close()


# ------------------------------------------------------------------ overlaps, ties
@pytest.mark.parametrize("prec", [S.PREC_MIXED, S.PREC_FP64])
def test_overlapping_pairs_match_dense_ewald(prec):
    # chains with pairs down to r = 0.5: the coefficient table for 1 <= r < 2 and
    # the closed form below r = 1 of the mixed-precision pass both run
    L, nb, nc, t = 20.0, 40, 3, 300
    p = S.default_params(nb=nb, nc=nc, L=L, flow_type=S.FLOW_SHEAR, edot=0.5, dt=1e-3, precision=prec)
    ctx = S.Context(p)
    pos = random_walk(nb, nc, L, 21, bond=1.0, min_sep=0.5)
    ctx.upload_positions(*xyz(pos))
    pw, cells = wrap_frame(ctx, t)
    d = pw[:, None, :] - pw[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))[np.triu_indices(len(pw), 1)]
    assert r.min() < 0.6 and np.count_nonzero(r < 1.0) > 20 and np.count_nonzero((r >= 1) & (r < 2)) > 100
    M = O.ewald_dense(cells[0], pw, 0.5)
    v = np.random.default_rng(11).normal(size=pw.size)
    e = rel(ctx.mobility_apply(v), M @ v)
    print(f"overlapping pairs (min r {r.min():.2f}), precision {prec}: M.F err {e:.2e}")
    assert e < 1e-5
    ctx.close()


def test_mc_acceptance_at_exact_tie():
    # A charge hops from x = 0 to x = 2 on a straight chain; the other two sit
    # at (1, +-10, 0), so every pair distance is unchanged and dE = 0 exactly in
    # any summation order.  With barrier 0 the move is accepted by 1 <= exp(0)
    # without a dice draw; the rest of the sweep follows the oracle.
    nb, nc, L = 5, 3, 40.0
    xs = np.array([-4.0, -2.0, 0.0, 2.0, 4.0])
    pos = np.concatenate([np.stack([xs, np.zeros(5), np.zeros(5)], 1),
                          np.stack([xs + 1.0, np.full(5, 10.0), np.zeros(5)], 1),
                          np.stack([xs + 1.0, np.full(5, -10.0), np.zeros(5)], 1)])
    track = np.array([2, 7, 12], np.int32)
    p = S.default_params(nb=nb, nc=nc, L=L, flow_type=S.FLOW_EQUILIBRIUM, edot=0.0, dt=1e-3, lambda_b=3.59,
                         lambda_d=5.0, barrier=0.0, barrier2=0.0)
    ctx = S.Context(p)
    ctx.upload_positions(*xyz(pos))
    ctx.remap(0)
    ctx.upload_charges([3], track)
    lat = O.lattice(0, L, 0.0, 1e-3, 1, 0)
    rx, ry, rz = xyz(np.stack(ctx.download_positions(), 1))
    u = np.array([0.9, 0.2, 0.7, 0.6, 0.4, 0.3, 0.8, 0.1, 0.55, 0.45], np.float32)
    cons, log = ctx.mc_sweep(S.MC_REFERENCE, u[None, :])
    mc = O.ChargeHop(nb, nc, track)
    st, cur, olog, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, 0.0, 0.0, u)
    assert olog[0, 2] == 3 and olog[0, 4] == 1 and olog[0, 5] == 1  # the tie: accepted on one draw
    assert cons[0] == cur and np.array_equal(log[:, 1], olog[:, 1]) and np.array_equal(log[:, 4:], olog[:, 4:])
    keep = olog[:, 3] >= 0
    assert np.array_equal(log[keep, 2:4], olog[keep, 2:4])
    assert np.array_equal(ctx.download_charges()[0], mc.track)
    ctx.close()
