"""GPU tests of the exactly detailed-balanced hop MC and its learned proposal,
through the C ABI, against the CPU oracle (oracle/smdlt_oracle.c orc_mh_*,
oracle/mlp_ref.py)."""
import numpy as np
import pytest
from helpers import O, S, random_walk

import mlp_ref as MLP
from test_gpu_parity import ctx_for, oracle_remap, xyz

pytestmark = pytest.mark.gpu


def mh_case(flow, nq, L=18.0, nb=30, nc=4, lambda_b=3.0, b=(0.0, 0.0), seed=12, t=40):
    ctx, p = ctx_for(flow, L, nb, nc, edot=0.3, lambda_d=5.0, lambda_b=lambda_b, barrier=b[0], barrier2=b[1])
    lat = O.lattice(flow, L, 0.3, 1e-3, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, seed))
    ctx.upload_positions(rx, ry, rz)
    ctx.remap(t)
    rx, ry, rz = oracle_remap(lat, t, rx, ry, rz)
    rng = np.random.default_rng(3)
    track = rng.choice(nb * nc, nq, replace=False).astype(np.int32)
    sign = np.where(np.arange(nq) % 2 == 0, 1, -1).astype(np.int32)
    ctx.upload_charges([nq], track, sign)
    op = O.mh_params(nb * nc, nb, hop_radius=3.0, feature_radius=4.0, floor_eps=0.05, lambda_b=lambda_b,
                     lambda_d=5.0, barrier=b[0], barrier2=b[1])
    qbead = np.zeros(nb * nc, np.int32)
    qbead[track] = sign
    return ctx, lat, (rx, ry, rz), track.copy(), sign, qbead, op


@pytest.mark.parametrize("flow,b", [(1, (0.0, 0.0)), (2, (0.3, 0.8)), (0, (1.0, 1.0))])
def test_mh_uniform_sequence_matches_oracle(flow, b):
    nq = 40
    ctx, lat, pos, track, sign, qbead, op = mh_case(flow, nq, b=b)
    rng = np.random.default_rng(7)
    for sweep in range(3):
        u = rng.random(3 * nq).astype(np.float32)
        cons, log = ctx.mc_sweep(S.MC_MH_UNIFORM, u[None, :])
        acc, olog, of = O.mh_sweep(lat, op, *pos, track, sign, qbead, None, u, features=True)
        assert cons[0] == 3 * nq
        assert np.array_equal(log, olog)
        tr, old = ctx.download_charges()
        assert np.array_equal(tr, track)
        f, fr, to, a = ctx.mc_features(0)
        assert np.array_equal(fr, olog[:, 1]) and np.array_equal(to, olog[:, 2]) and np.array_equal(a, olog[:, 4])
        assert np.allclose(f, of, rtol=1e-5, atol=1e-6)


def test_mh_learned_sequence_matches_oracle():
    nq = 40
    ctx, lat, pos, track, sign, qbead, op = mh_case(1, nq, b=(0.2, 0.5))
    w = MLP.init(np.random.default_rng(4), scale=2.0)
    ctx.mlp_set_weights(w)
    rng = np.random.default_rng(8)
    for sweep in range(3):
        u = rng.random(3 * nq).astype(np.float32)
        cons, log = ctx.mc_sweep(S.MC_MH_LEARNED, u[None, :])
        acc, olog, of = O.mh_sweep(lat, op, *pos, track, sign, qbead, w, u, features=True)
        assert np.array_equal(log, olog)
        f, fr, to, a = ctx.mc_features(0)
        assert np.allclose(f, of, rtol=1e-5, atol=1e-6)


def test_mlp_loss_grad_matches_oracle():
    rng = np.random.default_rng(9)
    w = MLP.init(rng)
    w[64 * 16:64 * 16 + 64] = rng.normal(0, 0.2, 64).astype(np.float32)
    w[-1] = 0.3
    X = rng.normal(size=(1000, 16)).astype(np.float32)  # 1000: the last 64-sample tile is partial
    y = (rng.random(1000) < 0.3).astype(np.float32)
    ctx, _ = ctx_for(1, 20.0, 10, 1)
    loss, g = ctx.mlp_loss_grad(w, X, y)
    wr = O.bf16_round(w).astype(np.float64)
    loss_ref, g_ref = MLP.loss_grad(wr, X.astype(np.float64), y.astype(np.float64))
    # BF16 operands in the 64x64 products: ~2^-8 relative per element
    assert loss == pytest.approx(loss_ref, rel=2e-3)
    o = 0
    for shape in MLP.SIZES:
        k = int(np.prod(shape))
        a, b = g[o:o + k], g_ref[o:o + k]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, (shape, o)
        o += k


def tau_int(x, c=5.0):
    """integrated autocorrelation time with Sokal's automatic window."""
    x = np.asarray(x, np.float64) - np.mean(x)
    n = len(x)
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf /= acf[0]
    tau = 1.0
    for W in range(1, n):
        tau = 1.0 + 2.0 * acf[1:W + 1].sum()
        if W >= c * tau:
            break
    return tau


def test_learned_proposal_trains_and_keeps_the_distribution():
    # 32 independent copies of one frozen conformation (device noise streams
    # differ by replica seed), hop radius 8: train on the uniform chains' moves,
    # then compare learned and uniform chains.  The learned proposal must accept
    # more, decorrelate the energy faster and leave the mean energy unchanged
    # (replica means are independent, so their spread gives the error bar).
    nq, R, nb, nc, L, hop = 40, 32, 30, 4, 18.0, 8.0
    rng = np.random.default_rng(3)
    track = rng.choice(nb * nc, nq, replace=False).astype(np.int32)
    sign = np.where(np.arange(nq) % 2 == 0, 1, -1).astype(np.int32)
    params = [S.default_params(nb=nb, nc=nc, L=L, flow_type=1, edot=0.3, dt=1e-3, lambda_d=5.0, lambda_b=4.0,
                               barrier=0.0, barrier2=0.0, seed=0x5EED0000 + r, hop_radius=hop) for r in range(R)]
    ctx = S.Context(params)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 12))
    ctx.upload_positions(np.tile(rx, R), np.tile(ry, R), np.tile(rz, R))
    ctx.remap(40)
    ctx.upload_charges([nq] * R, np.tile(track, R), np.tile(sign, R))
    F, Y = [], []
    for sweep in range(100):
        ctx.mc_sweep(S.MC_MH_UNIFORM)
        for r in range(R):
            f, fr, to, a = ctx.mc_features(r)
            keep = to >= 0
            F.append(f[keep])
            Y.append(a[keep])
    F = np.concatenate(F)
    Y = np.concatenate(Y).astype(np.float32)
    w, l0, l1 = ctx.mlp_train(F, Y, MLP.init(np.random.default_rng(10)), iters=800)
    assert l1 < 0.6 * l0
    ctx.mlp_set_weights(w)

    def chains(mode, sweeps=500, burn=100):
        ctx.upload_charges([nq] * R, np.tile(track, R), np.tile(sign, R))
        E = np.zeros((sweeps, R))
        acc = 0
        for s in range(sweeps):
            _, log = ctx.mc_sweep(mode)
            acc += log[:, 4].sum()
            E[s] = ctx.mc_energy()
        return E[burn:], acc / (sweeps * nq * R)

    Eu, au = chains(S.MC_MH_UNIFORM)
    El, al = chains(S.MC_MH_LEARNED)
    tu, tl = tau_int_multi(Eu), tau_int_multi(El)
    mu, ml = Eu.mean(0), El.mean(0)
    se = np.sqrt(mu.var(ddof=1) / R + ml.var(ddof=1) / R)
    print(f"acceptance uniform {au:.3f} learned {al:.3f}; tau_int(E) uniform {tu:.2f} learned {tl:.2f}; "
          f"<E> {mu.mean():.3f} {ml.mean():.3f} +- {se:.3f}")
    assert al > au + 0.05
    assert tl < tu
    assert abs(mu.mean() - ml.mean()) < 4 * se


def tau_int_multi(E, c=5.0):
    """tau_int of columns (independent replicas) from their averaged autocorrelation."""
    x = E - E.mean(0)
    n = len(x)
    f = np.fft.rfft(x, 2 * n, axis=0)
    acf = np.fft.irfft(f * np.conj(f), axis=0)[:n].mean(1)
    acf /= acf[0]
    tau = 1.0
    for W in range(1, n):
        tau = 1.0 + 2.0 * acf[1:W + 1].sum()
        if W >= c * tau:
            break
    return tau


@pytest.mark.parametrize("learned", [False, True])
def test_mh_many_rounds_match_oracle(learned):
    # 150 charges: three rounds of 64 attempts with the potential folded between them
    nq, nb, nc, L = 150, 60, 6, 26.0
    ctx, lat, pos, track, sign, qbead, op = mh_case(1, nq, L=L, nb=nb, nc=nc, b=(0.1, 0.4), seed=21)
    w = MLP.init(np.random.default_rng(9), scale=2.0) if learned else None
    if learned:
        ctx.mlp_set_weights(w)
    mode = S.MC_MH_LEARNED if learned else S.MC_MH_UNIFORM
    rng = np.random.default_rng(11)
    tot = 0
    for sweep in range(3):
        u = rng.random(3 * nq).astype(np.float32)
        cons, log = ctx.mc_sweep(mode, u[None, :])
        acc, olog, _ = O.mh_sweep(lat, op, *pos, track, sign, qbead, w, u)
        assert cons[0] == 3 * nq and np.array_equal(log, olog)
        assert np.array_equal(ctx.download_charges()[0], track)
        tot += acc
    assert tot > 40
    ctx.close()


def test_mh_stream_exhaustion_and_feature_switch():
    nq = 150
    ctx, lat, pos, track, sign, qbead, op = mh_case(1, nq, L=26.0, nb=60, nc=6, b=(0.1, 0.4), seed=22)
    u = np.random.default_rng(12).random(3 * 100 + 2).astype(np.float32)
    from test_gpu_parity import _raw_sweep
    code, cons, log = _raw_sweep(ctx, S.MC_MH_UNIFORM, u[None, :], nq)
    acc, olog, _ = O.mh_sweep(lat, op, *pos, track, sign, qbead, None, u[:300])
    assert code == -9 and cons[0] == 300
    assert np.array_equal(log[:100], olog[:100])
    assert np.array_equal(ctx.download_charges()[0], track)
    # without feature records the sweep decides the same moves
    ctx.mc_record_features(False)
    u2 = np.random.default_rng(13).random(3 * nq).astype(np.float32)
    cons, log = ctx.mc_sweep(S.MC_MH_UNIFORM, u2[None, :])
    acc, olog, _ = O.mh_sweep(lat, op, *pos, track, sign, qbead, None, u2)
    assert np.array_equal(log, olog)
    with pytest.raises(S.SmdltError):
        ctx.mc_features(0)
    ctx.mc_record_features(True)
    ctx.close()


def test_mh_large_replica_matches_oracle():
    # 40,000 beads: beyond the sizes whose occupancy stays in LDS
    nb, nc, L, nq = 400, 100, 110.0, 300
    ctx, p = ctx_for(1, L, nb, nc, edot=0.3, lambda_d=5.0, lambda_b=3.0, barrier=0.2, barrier2=0.5)
    lat = O.lattice(1, L, 0.3, 1e-3, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 23))
    ctx.upload_positions(rx, ry, rz)
    ctx.remap(5)
    rx, ry, rz = oracle_remap(lat, 5, rx, ry, rz)
    rng = np.random.default_rng(14)
    track = rng.choice(nb * nc, nq, replace=False).astype(np.int32)
    sign = np.where(np.arange(nq) % 2 == 0, 1, -1).astype(np.int32)
    ctx.upload_charges([nq], track, sign)
    op = O.mh_params(nb * nc, nb, hop_radius=3.0, feature_radius=4.0, floor_eps=0.05, lambda_b=3.0,
                     lambda_d=5.0, barrier=0.2, barrier2=0.5)
    qbead = np.zeros(nb * nc, np.int32)
    qbead[track] = sign
    track = track.copy()
    u = rng.random(3 * nq).astype(np.float32)
    cons, log = ctx.mc_sweep(S.MC_MH_UNIFORM, u[None, :])
    acc, olog, _ = O.mh_sweep(lat, op, rx, ry, rz, track, sign, qbead, None, u)
    print(f"40000-bead MH sweep: {acc} of {nq} accepted")
    assert np.array_equal(log, olog) and np.array_equal(ctx.download_charges()[0], track)
    ctx.close()
