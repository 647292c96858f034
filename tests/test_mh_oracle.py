"""CPU pins of the Metropolis-Hastings hop oracle and the proposal network restatement."""
import itertools

import numpy as np
import pytest
from helpers import O

import mlp_ref as MLP


def acceptance(lnrho, b):
    a, s = np.exp(-b), np.exp(0.5 * lnrho)
    return a * s / max(1.0, a * s, a / s)


def test_acceptance_is_detailed_balanced_and_reduces_to_reference_rule():
    rng = np.random.default_rng(0)
    for _ in range(200):
        lnrho, b = rng.normal(0, 4), rng.uniform(0, 3)
        assert 0 < acceptance(lnrho, b) <= 1
        assert acceptance(lnrho, b) / acceptance(-lnrho, b) == pytest.approx(np.exp(lnrho), rel=1e-12)
        if abs(lnrho) <= 2 * b:  # symmetric proposal: rho = exp(-dE); reference rule exp(-b - dE/2)
            assert acceptance(lnrho, b) == pytest.approx(np.exp(-b + 0.5 * lnrho), rel=1e-12)


def test_mlp_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    w = MLP.init(rng).astype(np.float64)
    w[64 * 16:64 * 16 + 64] = rng.normal(0, 0.1, 64)
    X = rng.normal(size=(40, 16))
    y = (rng.random(40) < 0.4).astype(np.float64)
    loss, g = MLP.loss_grad(w, X, y)
    idx = rng.choice(MLP.NPARAM, 60, replace=False)
    for i in idx:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (MLP.loss_only(wp, X, y) - MLP.loss_only(wm, X, y)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_mlp_forward_c_and_numpy_agree():
    rng = np.random.default_rng(2)
    w = MLP.init(rng)
    X = rng.normal(size=(5, 16))
    z, _ = MLP.forward(w, X)
    for k in range(5):
        assert O.mlp_forward(w, X[k]) == pytest.approx(z[k], rel=1e-12, abs=1e-12)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -3.14159, 1e-3], np.float32)  # 1+2^-8 ties to 1, 1+3*2^-8 to 1+2^-6
    r = O.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.015625)
    assert abs(r[3] - x[3]) <= abs(x[3]) * 2 ** -8 and abs(r[4] - x[4]) <= abs(x[4]) * 2 ** -8


def line_system(nb=8, spacing=2.05, L=60.0):
    """one straight chain along x, cubic box (flow 0)."""
    lat = O.lattice(0, L, 0.0, 1e-3, 1, 0)
    rx = np.arange(nb) * spacing - (nb - 1) * spacing / 2
    return lat, np.ascontiguousarray(rx), np.zeros(nb), np.zeros(nb)


def test_features_closed_form_two_charges():
    lat, rx, ry, rz = line_system()
    p = O.mh_params(8, 8, hop_radius=4.5, feature_radius=4.0, lambda_b=2.0, lambda_d=3.0)
    qbead = np.zeros(8, np.int32)
    qbead[3], qbead[5] = 1, -1
    f = O.mh_features(lat, p, rx, ry, rz, qbead, 3, 4, 1, ncand=3)
    k = 1 / 3.0
    e = lambda r: np.exp(-k * r) / r
    assert f[0] == pytest.approx(2.05) and f[1] == 1 and f[2] == pytest.approx(0.05)
    # partner at bead 5: 2.05 from bead 4 (< 4, so the pair counts) and 4.1 from bead 3
    assert f[3] == pytest.approx(2.0 * (1 * -1) * (e(2.05) - e(4.1)))
    assert f[4] == 1 and f[5] == 0 and f[6] == -1 and f[7] == 0
    assert f[8] == pytest.approx(0.1) and f[9] == pytest.approx(0.1)  # bead 5 within 4 of bead 4; bead 2 of bead 3
    assert f[10] == pytest.approx(-1.0) and f[11] == pytest.approx(-1.0)
    assert f[12] == 0 and f[13] == 0 and f[14] == 1 and f[15] == pytest.approx(0.3)


def exact_distribution(lat, p, rx, ry, rz, signs):
    nb = p.n
    states = [s for s in itertools.permutations(range(nb), len(signs))]
    E = []
    for s in states:
        q = np.zeros(nb, np.int32)
        q[list(s)] = signs
        E.append(O.coulomb_energy(lat, rx, ry, rz, q, p.lambda_b, p.lambda_d))
    pi = np.exp(-np.array(E))
    return states, pi / pi.sum()


@pytest.mark.parametrize("learned,barriers", [(False, (0.0, 0.0)), (True, (0.0, 0.0)), (True, (0.4, 1.1))])
def test_mh_chain_samples_boltzmann(learned, barriers):
    # 2 opposite charges on a 7-bead rod: 42 states, exact Boltzmann weights by
    # enumeration; the chain's visit frequencies over 1e6 attempts must match them
    # (a missing proposal ratio or barrier asymmetry shows up as a TV distance ~0.05+)
    lat, rx, ry, rz = line_system(nb=7, spacing=1.6)
    p = O.mh_params(7, 7, hop_radius=3.5, lambda_b=2.5, lambda_d=4.0, barrier=barriers[0], barrier2=barriers[1])
    signs = np.array([1, -1], np.int32)
    states, pi = exact_distribution(lat, p, rx, ry, rz, signs)
    index = {s: i for i, s in enumerate(states)}
    rng = np.random.default_rng(5)
    w = MLP.init(rng, scale=2.0) if learned else None
    track = np.array([0, 6], np.int32)
    qbead = np.zeros(7, np.int32)
    qbead[track] = signs
    counts = np.zeros(len(states))
    cur = list(track)
    for chunk in range(10 if not learned else 3):
        u = rng.random(3 * 100000).astype(np.float32)
        acc, log, _ = O.mh_sweep(lat, p, rx, ry, rz, track, signs, qbead, w, u)
        assert (log[:, 4] == 1).sum() == acc
        last = 0
        for a in np.nonzero(log[:, 4] == 1)[0]:
            counts[index[tuple(cur)]] += a - last
            cur[log[a, 0]] = log[a, 2]
            last = a
        counts[index[tuple(cur)]] += len(log) - last
        assert tuple(cur) == tuple(track)
    freq = counts / counts.sum()
    tv = 0.5 * np.abs(freq - pi).sum()
    print("TV distance", tv)
    assert tv < (0.012 if not learned else 0.02)
