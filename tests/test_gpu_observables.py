"""Output-step observables on the GPU against the oracle: the virial stress that
every step adds (bondforce/LJforce accumulators, S/main26.c:576-578,606-613,
662-664), the per-chain extension, centre of mass and gyration tensor
(calcExt/calcCOM/calcRg, S/main26.c:74-93,925-962,2391-2415) and the charge
displacement log of steps and hops (dxt, S/main26.c:815-838,2221-2233,2436-2446)."""
import numpy as np
import pytest
from helpers import O, S, random_walk
from test_gpu_parity import ctx_for, oracle_remap, xyz

pytestmark = pytest.mark.gpu


def setup(flow, fene, nb=40, nc=5, L=28.0, kappab=0.7, seed=7):
    kw = dict(spring_fene=fene, kappas=30.0 if fene else 200.0, qmax=3.0 if fene else 2.0, kappab=kappab,
              lj_ext_form=1 if flow == 2 else 0)
    ctx, p = ctx_for(flow, L, nb, nc, edot=0.2 if flow else 0.0, **kw)
    lat = O.lattice(flow, L, 0.2 if flow else 0.0, 1e-3, 1)
    pos = random_walk(nb, nc, L, seed, bond=2.4 if fene else 2.0)
    rx, ry, rz = xyz(pos)
    ctx.upload_positions(rx, ry, rz)
    return ctx, lat, kw, (rx, ry, rz)


def oracle_virial(lat, nb, nc, rx, ry, rz, kw, flow):
    st, _, tb = O.bond_force(lat, nb, nc, rx, ry, rz, kw["spring_fene"], kw["kappas"], kw["qmax"], kw["kappab"])
    assert st == 0
    nl, lst = O.verlet(lat, rx, ry, rz, 6.0)
    _, tl = O.lj_force(lat, rx, ry, rz, nl, lst, 5.03968 if flow == 2 else 25.0, 0.1, flow == 2)
    return tb + tl


@pytest.mark.parametrize("flow,fene", [(1, 0), (1, 1), (2, 0), (0, 0)])
def test_step_stress_equals_oracle_virial(flow, fene):
    nb, nc = 40, 5
    ctx, lat, kw, (rx, ry, rz) = setup(flow, fene, nb, nc)
    ctx.step(50)  # the step's forces are taken at the remapped positions of step 50
    tau = ctx.stress()[0]
    ref = oracle_virial(lat, nb, nc, *oracle_remap(lat, 50, rx, ry, rz), kw, flow)
    assert np.abs(tau - ref).max() <= 1e-10 * np.abs(ref).max()
    # read without reset keeps the sums; reset zeroes them after the read
    assert np.array_equal(ctx.stress(reset=True)[0], tau)
    assert np.array_equal(ctx.stress()[0], np.zeros(3))
    ctx.close()


def test_stress_graph_replay_equals_eager_steps():
    sums = []
    for graph in (False, True):
        ctx, lat, kw, _ = setup(1, 0, seed=9)
        if graph:
            ctx.run(10, 4, use_graph=True)
        else:
            for t in range(10, 14):
                ctx.step(t)
        sums.append(ctx.stress())
        ctx.close()
    assert np.array_equal(sums[0], sums[1])


def test_stress_of_batched_replicas():
    # two replicas with different chains: each replica's sums are its own
    specs = [(30, 4, 26.0, 1), (20, 6, 30.0, 2)]
    ps = [S.default_params(nb=nb, nc=nc, L=L, flow_type=fl, edot=0.2, dt=1e-3, kappas=200.0, qmax=2.0, kappab=0.5,
                           lj_ext_form=1 if fl == 2 else 0, seed=0x5EED0000 + r, pbc_compat=1)
          for r, (nb, nc, L, fl) in enumerate(specs)]
    ctx = S.Context(ps)
    pos = [random_walk(nb, nc, L, 20 + r) for r, (nb, nc, L, _) in enumerate(specs)]
    allp = np.concatenate(pos)
    ctx.upload_positions(*xyz(allp))
    ctx.step(30)
    tau = ctx.stress()
    kw = dict(spring_fene=0, kappas=200.0, qmax=2.0, kappab=0.5)
    for r, (nb, nc, L, fl) in enumerate(specs):
        lat = O.lattice(fl, L, 0.2, 1e-3, 1)
        ref = oracle_virial(lat, nb, nc, *oracle_remap(lat, 30, *xyz(pos[r])), kw, fl)
        assert np.abs(tau[r] - ref).max() <= 1e-10 * np.abs(ref).max(), r
    ctx.close()


@pytest.mark.parametrize("flow", [1, 2, 0])
def test_chain_observables_equal_oracle(flow):
    nb, nc = 40, 5
    ctx, lat, kw, (rx, ry, rz) = setup(flow, 0, nb, nc)
    ctx.remap(50)
    ext, com, rg = ctx.chain_observables()
    oe, oc, og = O.chain_obs(lat, nb, nc, *oracle_remap(lat, 50, rx, ry, rz))
    assert np.array_equal(ext, oe) and np.array_equal(com, oc) and np.array_equal(rg, og)
    ctx.close()


def test_chain_observables_of_batched_replicas():
    specs = [(30, 4, 26.0, 1), (20, 6, 30.0, 2), (50, 1, 40.0, 1)]
    ps = [S.default_params(nb=nb, nc=nc, L=L, flow_type=fl, edot=0.2, dt=1e-3, lj_ext_form=1 if fl == 2 else 0,
                           pbc_compat=1)
          for nb, nc, L, fl in specs]
    ctx = S.Context(ps)
    pos = [random_walk(nb, nc, L, 40 + r) for r, (nb, nc, L, _) in enumerate(specs)]
    ctx.upload_positions(*xyz(np.concatenate(pos)))
    ctx.remap(12)
    ext, com, rg = ctx.chain_observables()
    k = 0
    for r, (nb, nc, L, fl) in enumerate(specs):
        lat = O.lattice(fl, L, 0.2, 1e-3, 1)
        oe, oc, og = O.chain_obs(lat, nb, nc, *oracle_remap(lat, 12, *xyz(pos[r])))
        assert np.array_equal(ext[k:k + nc], oe) and np.array_equal(com[k:k + nc], oc), r
        assert np.array_equal(rg[k:k + nc], og), r
        k += nc
    ctx.close()


def c_round(x):
    """C's round(): halves away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def offsets_from_com(lat, nb, nc, rx, ry, rz):
    """com - p_j of every bead: the unwrapped chains of calcCOM (getNID walk)."""
    rel = np.zeros((nb * nc, 3))
    for k in range(nc):
        p = np.zeros((nb, 3))
        for j in range(nb - 1):
            p[j + 1] = p[j] + O.nid(lat, rx, ry, rz, k * nb + j, k * nb + j + 1)
        rel[k * nb:(k + 1) * nb] = p.mean(0) - p
    return rel


@pytest.mark.parametrize("flow", [1, 2])
def test_step_charge_log_matches_oracle(flow):
    # updateFD's log (S/main26.c:2436-2446; X/main14.c:853-876 for the flow part):
    # dt F + sqrt(2 dt) R of the tracked bead, plus the flow relative to the chain's
    # unwrapped centre of mass at the step start
    L, nb, nc, dt, edot = 24.0, 30, 3, 1e-3, 0.5
    ctx, p = ctx_for(flow, L, nb, nc, prec=S.PREC_FP64, edot=edot, dt=dt, kappas=200.0,
                     noise_mode=S.NOISE_FREE_DRAINING, lambda_b=3.59, lambda_d=5.0,
                     lj_ext_form=1 if flow == 2 else 0)
    lat = O.lattice(flow, L, edot, dt, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 9))
    ctx.upload_positions(rx, ry, rz)
    track = np.array([3, 17, 31, 44, 60, 89], np.int32)
    ctx.upload_charges([len(track)], track)
    z = np.random.default_rng(8).normal(size=3 * nb * nc)
    t = 10
    ctx.step(t, z)
    log = ctx.charge_displacements()
    wx, wy, wz = oracle_remap(lat, t, rx, ry, rz)
    nl, lst = O.verlet(lat, wx, wy, wz, 6.0)
    _, fb, _ = O.bond_force(lat, nb, nc, wx, wy, wz, 0, 200.0, 2.0)
    fl, _ = O.lj_force(lat, wx, wy, wz, nl, lst, 5.03968 if flow == 2 else 25.0, 0.1, flow == 2)
    q = np.zeros(nb * nc, np.int32)
    q[track] = 1
    F = fb + fl + O.coulomb_force(lat, wx, wy, wz, q, 3.59, 5.0)  # lab frame
    d = dt * F + np.sqrt(2 * dt) * z.reshape(-1, 3)
    rel = offsets_from_com(lat, nb, nc, wx, wy, wz)
    exp = d[track].copy()
    if flow == 2:
        exp[:, 0] += dt * edot * rel[track, 0]
        exp[:, 1] -= dt * edot * rel[track, 1]
    else:
        exp[:, 0] += dt * edot * rel[track, 1]
    assert np.abs(log - exp).max() <= 1e-10 * np.abs(exp).max()
    assert np.array_equal(ctx.charge_displacements(reset=True), log)
    assert not ctx.charge_displacements().any()
    ctx.close()


def test_mc_hops_add_their_vectors_to_the_charge_log():
    # chargeHop adds new bead - old bead, reduced by the cell without the shear
    # shift, for every accepted hop (S/main26.c:2221-2233)
    L, nb, nc, nq = 22.0, 30, 4, 24
    ctx, p = ctx_for(1, L, nb, nc, edot=0.3, lambda_d=5.0, lambda_b=3.59, barrier=0.2, barrier2=0.5)
    ctx.upload_positions(*xyz(random_walk(nb, nc, L, 12)))
    lat_row = ctx.update_lattice(40)[0]
    ctx.remap(40)
    track = np.random.default_rng(3).choice(nb * nc, nq, replace=False).astype(np.int32)
    ctx.upload_charges([nq], track)
    rng = O.Ran1(830062920)
    x, y, zc = ctx.download_positions()
    L1xp, L2xp, L2yp = lat_row[4], lat_row[5], lat_row[6]
    total = np.zeros((nq, 3))
    accepted = 0
    for sweep in range(3):
        _, log = ctx.mc_sweep(S.MC_REFERENCE, rng.uniforms(3 * nq)[None, :])
        for k, c, tgt, kind, acc, nd in log:
            if acc:
                accepted += 1
                dx, dy, dz = x[tgt] - x[c], y[tgt] - y[c], zc[tgt] - zc[c]
                dx -= L2xp * c_round(dy / L2yp)
                dy -= L2yp * c_round(dy / L2yp)
                dx -= L1xp * c_round(dx / L1xp)
                dz -= L * c_round(dz / L)
                total[k] += (dx, dy, dz)
    assert accepted > 0
    assert np.abs(ctx.charge_displacements() - total).max() <= 1e-12 * max(1.0, np.abs(total).max())
    ctx.close()
