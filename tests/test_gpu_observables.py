"""Output-step observables on the GPU against the oracle: the virial stress that
every step adds (bondforce/LJforce accumulators, S/main26.c:576-578,606-613,
662-664) and the per-chain extension, centre of mass and gyration tensor
(calcExt/calcCOM/calcRg, S/main26.c:74-93,925-962,2391-2415)."""
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
                           lj_ext_form=1 if fl == 2 else 0, seed=0x5EED0000 + r)
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
    ps = [S.default_params(nb=nb, nc=nc, L=L, flow_type=fl, edot=0.2, dt=1e-3, lj_ext_form=1 if fl == 2 else 0)
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
