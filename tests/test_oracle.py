"""Pins of the CPU oracle (the checker) against closed forms, invariants,
brute force and values of the reference's calcHI computed by an independent restatement
(the self term 0.95272918 at L = 60, 0.94307920 with its self-image loop, and the
overlap-branch errors below).
These run without a GPU."""
import numpy as np
import pytest
from helpers import O, random_walk

import spme_ref as P


def xyz(pos):
    return [np.ascontiguousarray(pos[:, k]) for k in range(3)]


# ----------------------------------------------------------------- RNG (ran1/gasdev)
def test_ran1_lcg_state_is_schrage_multiplication():
    # idum_{k+1} = 40014 * idum_k mod 2147483563 (L'Ecuyer 1988), computed here
    # with exact integers instead of Schrage's decomposition.
    r = O.Ran1(830062920)
    x = 830062920
    for _ in range(1000):
        r.uniforms(1)
        x = (40014 * x) % 2147483563
        assert r.s.idum == x


def test_ran1_positive_seed_quirks():
    # positive seeds skip the shuffle-table fill, so the first two draws are
    # equal, and idum2 never advances (S/Initialization.c:609-627)
    u = O.Ran1(830062920).uniforms(3)
    assert u[0] == u[1]
    r = O.Ran1(981272199)
    r.uniforms(500)
    assert r.s.idum2 == 123456789


def test_ran1_range_and_moments():
    u = O.Ran1(12345).uniforms(200000).astype(np.float64)
    assert u.min() > 0 and u.max() <= 1 - 1.2e-7 + 1e-12
    assert abs(u.mean() - 0.5) < 0.003
    g = O.Ran1(12345).gauss(200000)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1) < 0.01


def test_ran1_negative_seed_fills_table():
    r = O.Ran1(-7)
    u = r.uniforms(2)
    assert u[0] != u[1]
    assert r.s.idum2 == 7


# ----------------------------------------------------------------- lattice / remap
def test_shear_lattice_is_deforming_brick():
    L, edot, dt = 30.0, 0.05, 1e-3
    lat = O.lattice(1, L, edot, dt)
    assert lat.tp == int(0.9624 / edot / dt)
    for t in (0, 17, lat.tp - 1, lat.tp + 5):
        O.lattice_update(lat, t)
        g = edot * (t % lat.tp) * dt
        assert lat.L1xp == pytest.approx(L, rel=1e-14) and abs(lat.L1yp) < 1e-12
        assert lat.L2xp == pytest.approx(g * L, abs=1e-12)
        assert lat.L2yp == pytest.approx(L, rel=1e-14)
        assert lat.delx == pytest.approx(g, abs=1e-15)


def test_krbc_lattice_returns_after_one_period():
    # Kraynik-Reinelt: after the Hencky strain ln((3+sqrt5)/2) the deformed
    # lattice is the initial one in a new integer basis.  The reference uses the
    # truncated EP=0.9624 and THETA0=31.7 deg, so equivalence holds to ~1e-4.
    L = 20.0
    edot, dt = 0.1, 1e-3
    lat = O.lattice(2, L, edot, dt)
    A0 = np.array([[lat.L1[0], lat.L2[0]], [lat.L1[1], lat.L2[1]]])
    pts = np.array(lat.point0)
    f = np.exp(edot * lat.tp * dt)
    A1 = (pts[2] - pts[1]) * [f, 1 / f], (pts[0] - pts[1]) * [f, 1 / f]
    A1 = np.array(A1).T
    Mint = np.linalg.solve(A0, A1)
    assert np.abs(Mint - np.round(Mint)).max() < 2e-3
    assert abs(np.linalg.det(np.round(Mint))) == pytest.approx(1.0)


@pytest.mark.parametrize("flow,compat", [(1, 1), (1, 0), (2, 1), (0, 1)])
def test_apply_pbc_wraps_into_cell(flow, compat):
    L = 25.0
    lat = O.lattice(flow, L, 0.2, 1e-3, compat, 2345)
    rng = np.random.default_rng(3)
    p = rng.uniform(-3 * L, 3 * L, (500, 3))
    rx, ry, rz = O.apply_pbc(lat, *xyz(p))
    assert np.all(np.abs(ry) <= lat.L2yp / 2 + 1e-9)
    assert np.all(np.abs(rz) <= L / 2 + 1e-9)
    assert np.all(np.abs(rx - ry * lat.L2xp / lat.L2yp) <= lat.L1xp / 2 + 1e-9)
    # idempotent
    rx2, ry2, rz2 = O.apply_pbc(lat, rx, ry, rz)
    assert np.allclose(rx2, rx) and np.allclose(ry2, ry) and np.allclose(rz2, rz)


@pytest.mark.parametrize("flow", [1, 2])
def test_min_image_lattice_shift_invariant(flow):
    # without the delx quirk getNID only depends on positions modulo the lattice
    L = 18.0
    lat = O.lattice(flow, L, 0.3, 1e-3, 0, 777)
    rng = np.random.default_rng(5)
    p = rng.uniform(-L / 2, L / 2, (40, 3))
    rx, ry, rz = xyz(p)
    a1 = np.array([lat.L1xp, 0.0, 0.0])
    a2 = np.array([lat.L2xp, lat.L2yp, 0.0])
    for i in range(0, 40, 7):
        for j in range(1, 40, 5):
            d0 = O.nid(lat, rx, ry, rz, i, j)
            q = p.copy()
            q[j] += 2 * a1 - a2 + np.array([0, 0, 3 * L])
            d1 = O.nid(lat, *xyz(q), i, j)
            assert np.allclose(d0, d1, atol=1e-9)


def test_min_image_is_true_minimum_within_half_width():
    L = 20.0
    lat = O.lattice(1, L, 0.5, 1e-3, 0, 300)
    cell = (lat.L1xp, lat.L2xp, lat.L2yp, L)
    rng = np.random.default_rng(9)
    p = rng.uniform(-L / 2, L / 2, (60, 3))
    rx, ry, rz = xyz(p)
    for i in range(60):
        for j in range(60):
            if i == j:
                continue
            d = O.nid(lat, rx, ry, rz, i, j)
            m = P.min_image(p[i] - p[j], cell)
            if np.linalg.norm(m) < 0.4 * L:
                assert np.allclose(d, m, atol=1e-9)


def test_krbc_rotation_round_trip_drift_is_ulps():
    lat = O.lattice(2, 30.0, 0.1, 1e-3, 1, 4000)
    rng = np.random.default_rng(2)
    rx, ry = rng.uniform(-15, 15, 200), rng.uniform(-15, 15, 200)
    a, b = O.rotate(lat, rx, ry, 0)
    c, d = O.rotate(lat, a, b, 1)
    assert np.abs(c - rx).max() < 1e-13 and np.abs(d - ry).max() < 1e-13


# ----------------------------------------------------------------- neighbour list
@pytest.mark.parametrize("flow,compat", [(1, 1), (1, 0), (2, 1)])
def test_verlet_matches_brute_force(flow, compat):
    L = 22.0
    pos = random_walk(30, 4, L, 11)
    lat = O.lattice(flow, L, 0.4, 1e-3, compat, 1500)
    rx, ry, rz = O.apply_pbc(lat, *xyz(pos))
    nl, lst = O.verlet(lat, rx, ry, rz, 6.0)
    n = len(rx)
    for i in range(n):
        expect = [j for j in range(i + 1, n) if np.linalg.norm(O.nid(lat, rx, ry, rz, i, j)) < 6.0]
        assert list(lst[i, :nl[i]]) == expect
    # with the exact wrap the list is the true minimum-image neighbour set
    if not compat:
        cell = (lat.L1xp, lat.L2xp, lat.L2yp, L)
        p = np.stack([rx, ry, rz], 1)
        for i in range(0, n, 9):
            d = np.linalg.norm(P.min_image(p[i + 1:] - p[i], cell), axis=1)
            assert sorted((np.nonzero(d < 6.0)[0] + i + 1).tolist()) == list(lst[i, :nl[i]])


# ----------------------------------------------------------------- forces
def test_two_bead_forces_closed_form():
    lat = O.lattice(0, 50.0, 0.0, 1e-3)
    r = 2.3
    rx, ry, rz = np.array([0.0, r]), np.zeros(2), np.zeros(2)
    st, fb, tau = O.bond_force(lat, 2, 1, rx, ry, rz, 0, 200.0, 2.0)
    assert st == 0
    assert fb[0, 0] == pytest.approx(200.0 * (r - 2.0))  # Hookean pulls together
    st, ff, _ = O.bond_force(lat, 2, 1, rx, ry, rz, 1, 30.0, 3.0)
    assert ff[0, 0] == pytest.approx(30.0 * r / (1 - r * r / 9.0))
    st, _, _ = O.bond_force(lat, 2, 1, np.array([0.0, 3.5]), ry, rz, 1, 30.0, 3.0)
    assert st == -1  # FENE overstretch (S/main26.c:553-558)
    nl, lst = O.verlet(lat, rx, ry, rz, 6.0)
    fl, _ = O.lj_force(lat, rx, ry, rz, nl, lst, 25.0, 0.1, 0)
    s6 = (4.0 / r ** 2) ** 3
    assert fl[1, 0] == pytest.approx(12 * 0.1 / r ** 2 * (s6 * s6 - s6) * r)
    q = np.array([1, 1], np.int32)
    fc = O.coulomb_force(lat, rx, ry, rz, q, 3.59, 5.0)
    assert fc[1, 0] == pytest.approx(3.59 * np.exp(-r / 5) * (1 + r / 5) / r ** 2)
    assert O.coulomb_energy(lat, rx, ry, rz, q, 3.59, 5.0) == pytest.approx(3.59 * np.exp(-r / 5) / r)


def test_virial_equals_clausius_sum():
    # pair forces f_i = c d, f_j = -c d with d = r_i - r_j give sum_pairs c d d =
    # sum_i r_i F_i, so the accumulators -2 sum_pairs d c d equal -2 sum_i r_i F_i
    # for an unwrapped chain that meets no image (springs and LJ; bending adds
    # its own quirky terms, S/main26.c:606-613)
    L = 400.0
    pos = random_walk(30, 2, L, 5)
    p = pos - pos.mean(0)
    p = np.concatenate([np.cumsum(np.vstack([p[:1] * 0, np.diff(p[:30], axis=0)]), 0),
                        np.cumsum(np.vstack([p[30:31] * 0 + 3.0, np.diff(p[30:], axis=0)]), 0)])
    lat = O.lattice(0, L, 0.0, 1e-3)
    rx, ry, rz = xyz(p)
    for fene, kappas, qmax in ((0, 200.0, 2.0), (1, 30.0, 3.0)):
        st, fb, tb = O.bond_force(lat, 30, 2, rx, ry, rz, fene, kappas, qmax)
        assert st == 0
        nl, lst = O.verlet(lat, rx, ry, rz, 6.0)
        fl, tl = O.lj_force(lat, rx, ry, rz, nl, lst, 25.0, 0.1, 0)
        for f, tau in ((fb, tb), (fl, tl)):
            ref = -2.0 * np.array([(p[:, 0] * f[:, 0]).sum(), (p[:, 1] * f[:, 1]).sum(), (p[:, 0] * f[:, 1]).sum()])
            assert np.abs(tau - ref).max() <= 1e-9 * np.abs(ref).max()
        assert np.abs(tl).max() > 0 and np.abs(tb).max() > 0


def test_chain_observables_closed_forms_and_wrap_invariance():
    # a straight rod of nb beads spaced b along x: extension (nb-1) b, Rg_xx =
    # b^2 (nb^2 - 1)/12, centre of mass at -(nb-1) b/2 (the unwrapping adds
    # getNID(j, j+1) = r_j - r_{j+1}, so chains come out mirrored)
    lat = O.lattice(1, 40.0, 0.3, 1e-3, 1, 500)
    nb, b = 12, 1.9
    x = np.arange(nb) * b - 10.0
    ext, com, rg = O.chain_obs(lat, nb, 1, x, np.full(nb, 0.4), np.full(nb, -3.0))
    assert ext[0] == pytest.approx((nb - 1) * b)
    assert com[0] == pytest.approx([-(nb - 1) * b / 2, 0.0, 0.0])
    assert rg[0] == pytest.approx([b * b * (nb * nb - 1) / 12, 0, 0, 0, 0, 0], abs=1e-12)
    # random chains: equal to the moments of the unwrapped coordinates, and unchanged
    # when beads are moved by lattice vectors (the minimum image undoes the wrap)
    L, nb, nc = 30.0, 25, 3
    pos = random_walk(nb, nc, L, 8)
    unw = np.concatenate([np.cumsum(np.vstack([np.zeros((1, 3)), np.diff(pos[k * nb:(k + 1) * nb], axis=0)
                                               - L * np.round(np.diff(pos[k * nb:(k + 1) * nb], axis=0) / L)]), 0)
                          for k in range(nc)])
    lat = O.lattice(0, L, 0.0, 1e-3)
    ext, com, rg = O.chain_obs(lat, nb, nc, *xyz(pos))
    for k in range(nc):
        u = -unw[k * nb:(k + 1) * nb]
        assert ext[k] == pytest.approx(u[:, 0].max() - u[:, 0].min(), rel=1e-12)
        assert com[k] == pytest.approx(u.mean(0), abs=1e-10)
        c = (u - u.mean(0)).T @ (u - u.mean(0)) / nb
        assert rg[k] == pytest.approx([c[0, 0], c[0, 1], c[0, 2], c[1, 1], c[1, 2], c[2, 2]], abs=1e-10)
    shifted = pos + L * np.random.default_rng(3).integers(-2, 3, size=pos.shape)
    e2, c2, g2 = O.chain_obs(lat, nb, nc, *xyz(shifted))
    assert np.allclose(e2, ext, rtol=1e-12) and np.allclose(c2, com, atol=1e-10) and np.allclose(g2, rg, atol=1e-10)


def test_forces_sum_to_zero():
    L = 30.0
    pos = random_walk(25, 3, L, 2)
    lat = O.lattice(1, L, 0.1, 1e-3, 0, 100)
    rx, ry, rz = O.apply_pbc(lat, *xyz(pos))
    st, fb, _ = O.bond_force(lat, 25, 3, rx, ry, rz, 0, 200.0, 2.0, kappab=1.5)
    nl, lst = O.verlet(lat, rx, ry, rz, 6.0)
    fl, _ = O.lj_force(lat, rx, ry, rz, nl, lst, 25.0, 0.1, 0)
    q = np.zeros(75, np.int32)
    q[::3] = 1
    q[1::7] = -1
    fc = O.coulomb_force(lat, rx, ry, rz, q, 3.59, 5.0)
    for f in (fb, fl, fc):
        assert np.abs(f.sum(0)).max() < 1e-9 * max(1.0, np.abs(f).max())


# ----------------------------------------------------------------- Ewald RPY
@pytest.mark.parametrize("L", [12.0, 20.0, 60.0])
def test_hasimoto_self_mobility(L):
    cell = O.Cell(L, 0, 0, L, L)
    for a in (3.0 / L, 6.0 / L):
        M = O.ewald_dense(cell, np.zeros(3), a)
        h = 1 - 2.8372974795 / L + 4.1887902048 / L ** 3
        assert M[0, 0] == pytest.approx(h, abs=30.0 / L ** 6 + 1e-9)
        assert np.allclose(M, M[0, 0] * np.eye(3), atol=1e-12)


def test_ewald_alpha_invariant_on_sheared_lattice():
    L = 15.0
    cell = O.Cell(L, 0, 0.37 * L, L, L)
    pos = np.random.default_rng(1).uniform(-L / 2, L / 2, (6, 3))
    Ms = [O.ewald_dense(cell, pos, a) for a in (0.2, 0.4, 0.8)]
    assert np.abs(Ms[0] - Ms[1]).max() < 1e-11
    assert np.abs(Ms[1] - Ms[2]).max() < 1e-11


def test_ewald_symmetric_positive_definite_with_overlaps():
    L = 14.0
    pos = random_walk(20, 2, L, 4, bond=1.5, min_sep=1.0)
    M = O.ewald_dense(O.Cell(L, 0, 0.2 * L, L, L), pos, 0.5)
    assert np.abs(M - M.T).max() < 1e-12
    assert np.linalg.eigvalsh(M).min() > 0


def test_ewald_large_box_approaches_free_rpy():
    pos = np.array([[0.0, 0, 0], [2.5, 0.7, -0.4], [1.0, 1.3, 0.2]])
    Mf = O.rpy_free_dense(pos)
    for L in (200.0, 400.0):
        Me = O.ewald_dense(O.Cell(L, 0, 0, L, L), pos, 6.0 / L)
        dev = np.abs(Me - Mf).max()
        assert dev < 3.2 / L and dev > 2.0 / L  # the leading periodic correction is -2.837/L


def test_overlap_branch_matches_independent_measurements():
    # relative error of the reference's r<=2 I-coefficient at alpha=4/60
    for r, e in ((1.0, 2.67e-2), (1.5, 4.13e-3), (1.9, 2.14e-3), (1.99, 3.21e-4)):
        te, _ = O.rpy_real_coeffs(r, 4 / 60, 0)
        tr, _ = O.rpy_real_coeffs(r, 4 / 60, 1)
        assert abs(tr - te) / abs(te) == pytest.approx(e, rel=0.01)
    # and the exact branch is continuous at r = 2 with the far form
    a, b = O.rpy_real_coeffs(2.0 - 1e-9, 0.3, 0)
    c, d = O.rpy_real_coeffs(2.0 + 1e-9, 0.3, 0)
    assert abs(a - c) < 1e-8 and abs(b - d) < 1e-8


def test_calchi_self_term_independent_values():
    L = 60.0
    lat = O.lattice(1, L, 0.01, 1e-3, 1, 100)  # t%tp <= tp/2: rce = L
    b = O.bins(lat, 2.0)
    pos = np.array([[0.0, 0, 0], [3.0, 1.0, 0.5]])
    M, far, kc = O.calchi_ref(lat, *xyz(pos), 100, 8, b, None, L ** 3)
    assert M[0, 0] == pytest.approx(0.95272918, abs=2e-8)
    # second half of a period with an undeformed lattice: rce = 2L, alpha = 2/L,
    # the self-image loop with && and dy*dy gives 0.94307920 (the reference's calcHI
    # evaluated by an independent restatement)
    lat.treal = lat.tp // 2 + 1
    lat.L2[0] = 0.0
    lat.L2xp = 0.0
    lat.delx = 0.0
    M2, _, _ = O.calchi_ref(lat, *xyz(pos), 0, 8, b, None, L ** 3)
    assert M2[0, 0] == pytest.approx(0.94307920, abs=2e-8)


def test_calchi_near_pairs_match_exact_ewald_in_declared_regime():
    # t%tp <= tp/2, all pairs closer than 12a, none overlapping: the reference's
    # matrix equals the exact Ewald sum up to its k-cut (~1e-6 relative)
    L = 13.0
    rng = np.random.default_rng(8)
    lat = O.lattice(1, L, 0.05, 1e-3, 1, 2000)
    cell = (lat.L1xp, lat.L2xp, lat.L2yp, L)
    pos = []
    while len(pos) < 12:  # random sequential addition, no pair closer than 2.2
        q = rng.uniform(-L / 2, L / 2, 3)
        if all(np.linalg.norm(P.min_image(q - x, cell)) > 2.2 for x in pos):
            pos.append(q)
    rx, ry, rz = O.apply_pbc(lat, *xyz(np.array(pos)))
    p = np.stack([rx, ry, rz], 1)
    b = O.bins(lat, 2.0)
    M, far, kc = O.calchi_ref(lat, rx, ry, rz, 2000, 8, b, None, L ** 3)
    Me = O.ewald_dense(O.cell_of(lat), p, 0.4)
    off = ~np.kron(np.eye(len(p), dtype=bool), np.ones((3, 3), bool))
    assert np.abs(M - Me)[off].max() / np.abs(Me[off]).max() < 2e-5


# ----------------------------------------------------------------- ewaldBin / CATEA
def _separated(lat, L, n, seed, sep=2.2):
    """random sequential addition with no pair closer than sep, wrapped by applyPBC"""
    rng = np.random.default_rng(seed)
    cell = (lat.L1xp, lat.L2xp, lat.L2yp, L)
    pos = []
    while len(pos) < n:
        q = rng.uniform(-L / 2, L / 2, 3)
        if all(np.linalg.norm(P.min_image(q - x, cell)) > sep for x in pos):
            pos.append(q)
    return O.apply_pbc(lat, *xyz(np.array(pos)))


def test_ewald_bin_self_block_is_the_ewald_diagonal():
    # first half of a period, rce = L: the images the reference's && drops lie at
    # r >= L = rce (erfc(3.5) ~ 7e-7), so Dtr is the Ewald self-mobility truncated
    # at ewaldBin's split (alpha = 3.5/L, k-cut 2*3.5^2/L): the dense Ewald sum with
    # the same split and cuts (its k range edge differs from the reference's by
    # ~1e-6), and within the M_HI = 3.5 truncation (~3e-5) of the converged value,
    # 0.95272918 at L = 60 (Hasimoto 1 - 2.837297/L + 4.18879/L^3)
    L = 60.0
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 1.0, 0.5]])
    for flow, t in ((0, 0), (1, 100)):
        lat = O.lattice(flow, L, 0.01 if flow else 0.0, 1e-3, 1, t)
        eps, rs, dtr, kc = O.ewald_bin(lat, *xyz(pos), 2, L ** 3)
        Mt = O.ewald_dense(O.cell_of(lat), pos, 3.5 / L, rcut=L, kcut=2 * 3.5 ** 2 / L)
        assert np.abs(dtr - np.diag(Mt)[:3]).max() < 2e-6
        assert np.abs(dtr - 0.95272918).max() < 5e-5
        assert abs(0.95272918 - (1 - 2.837297 / L + 4.18879 / L ** 3)) < 3e-6


@pytest.mark.parametrize("flow,t", [(1, 2000), (1, 15000), (2, 700)])
def test_ewald_bin_pair_sums_match_exact_ewald(flow, t):
    # no pair closer than 2, so the pair blocks are the alpha-independent Ewald
    # sum.  Driven to convergence (M_HI = 6: erfc(6) and the k-cut are ~1e-16) the
    # restated sums (reference k range and its bounds, images within rce, raw
    # differences, both halves of the period and the KRBC lattice in the lab
    # frame) equal eps and rowsum rebuilt from the dense Ewald matrix with
    # ewaldBin's own Dtr; at the reference's M_HI = 3.5 they carry its ~5e-4
    # truncation.
    L, n = 13.0, 12
    lat = O.lattice(flow, L, 0.05, 1e-3, 1, t)
    p = np.stack(_separated(lat, L, n, 8), 1)  # in the cell frame (a1 along x)
    c, s_ = lat.ct, lat.st
    Rm = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1.0]]) if flow == 2 else np.eye(3)  # cell frame -> lab
    R9 = np.kron(np.eye(n), Rm)
    Me = R9 @ O.ewald_dense(O.cell_of(lat), p, 0.4) @ R9.T
    off = ~np.kron(np.eye(n, dtype=bool), np.ones((3, 3), bool))
    for m_hi, tol in ((6.0, 1e-11), (3.5, 2e-3)):
        eps, rs, dtr, kc = O.ewald_bin(lat, *xyz(p @ Rm.T), 4, L ** 3, m_hi=m_hi)
        W = np.where(off, Me, 0.0) / np.repeat(dtr[None, :], n, 0).reshape(-1)[:, None]
        assert abs(eps - W.sum()) / np.abs(W).sum() < tol
        rs_dense = (W ** 2).sum(1)
        assert np.abs(rs - rs_dense).max() / rs_dense.max() < tol


@pytest.mark.parametrize("flow", [1, 2])
def test_far_table_equals_converged_ewald_pair_blocks(flow):
    # gridHI's split (M_HI = 6) is converged, so away from overlap every table entry
    # is the exact Ewald pair block of the displacement at the lattice of its
    # strain phase (t = nt*tp/nt_avg), where the reference's image range holds every
    # image within rce.  Shear: all phases, both halves of the period.  KRBC: the
    # first half; later phases are so skewed that the reference's contiguous
    # ky/kz ranges drop in-cut wave vectors (1e-8..4e-5), which the restatement keeps.
    L, nt_avg = 12.0, 4
    lat = O.lattice(flow, L, 0.05, 1e-3, 1, 0)
    b = O.bins(lat, 2.0)
    tab = O.far_table(lat, nt_avg, b, L ** 3)
    nb = b.nbx * b.nby * b.nbz
    assert tab.shape == (nt_avg, nb, 3, 3)
    idx = np.arange(nb)
    bx, rem = idx // (b.nby * b.nbz), idx % (b.nby * b.nbz)
    dr = np.stack([(bx - b.nhx) * b.bin_size_x, (rem // b.nbz - b.nhy) * b.bin_size_y,
                   (rem % b.nbz - b.nhz) * b.bin_size_z], 1)
    origin = b.nhx * b.nby * b.nbz + b.nhy * b.nbz + b.nhz
    dr[origin] = 0.6
    checked = 0
    g = np.arange(-6, 7)
    ns = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    for nt in range(nt_avg if flow == 1 else 2):
        latn = O.lattice(flow, L, 0.05, 1e-3, 1, nt * lat.tp // nt_avg)
        c, s_ = latn.ct, latn.st
        Rm = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1.0]]) if flow == 2 else np.eye(3)  # cell frame -> lab
        cell = O.cell_of(latn)
        cellt = (cell.a1x, cell.a2x, cell.a2y, cell.Lz)
        first = latn.treal <= latn.tp // 2
        rce, nmax = (latn.L2yp, 1) if first else (2 * latn.L2yp, 2)
        shifts = ns[:, :1] * np.array([latn.L1[0], latn.L1[1], 0]) + ns[:, 1:2] * np.array(
            [latn.L2[0], latn.L2[1], 0]) + ns[:, 2:] * np.array([0, 0, L])
        for i in range(0, nb, 3):
            d = dr[i] @ Rm  # the displacement in the cell frame
            if np.linalg.norm(P.min_image(d, cellt)) < 2.2:
                continue
            # only where the reference's image range holds every image within rce (on
            # skewed KRBC phases and for displacements beyond the box it does not)
            inside = np.linalg.norm(dr[i] + shifts, axis=1) < rce
            if np.abs(ns[inside]).max(initial=0) > nmax:
                continue
            Me = O.ewald_dense(cell, np.array([[0.0, 0, 0], d]), 0.5)
            ref = Rm @ Me[0:3, 3:6] @ Rm.T
            assert np.abs(tab[nt, i] - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())
            checked += 1
    assert checked > 100


def test_catea_beta_is_the_small_root_and_dilute_limit():
    # beta solves X b^2 - 2b + 1 = 0 (X = 3N(e^2 - e), e = eps/(9N^2 - 3N)) with the
    # smaller root, and tends to 1/2 as the coupling vanishes; C decreases with rowsum
    L, n = 13.0, 12
    lat = O.lattice(1, L, 0.05, 1e-3, 1, 2000)
    rx, ry, rz = _separated(lat, L, n, 9)
    eps, rs, dtr, kc = O.ewald_bin(lat, rx, ry, rz, 4, L ** 3)
    beta, Cv = O.catea(n, eps, rs)
    e = eps / (9 * n * n - 3 * n)
    X = 3 * n * (e * e - e)
    assert abs(X * beta * beta - 2 * beta + 1) < 1e-12
    assert beta < 1.0 / X * (1 + np.sqrt(1 - X)) if X > 0 else beta > 0
    order = np.argsort(rs)
    assert np.all(np.diff(Cv[order]) <= 1e-15) and Cv.max() <= 1.0
    lat2 = O.lattice(0, 400.0, 0.0, 1e-3, 1, 0)
    eps2, rs2, _, _ = O.ewald_bin(lat2, np.array([0.0, 60.0]), np.zeros(2), np.zeros(2), 1, 400.0 ** 3)
    b2, _ = O.catea(2, eps2, rs2)
    assert abs(b2 - 0.5) < 0.01


# ----------------------------------------------------------------- Brownian helpers
def test_tea_reduces_to_mobility_and_diagonal():
    L = 16.0
    pos = random_walk(10, 2, L, 1)
    M = O.ewald_dense(O.Cell(L, 0, 0, L, L), pos, 0.5)
    z = np.random.default_rng(0).normal(size=60)
    one = np.ones(60)
    assert np.allclose(O.tea_apply(M, 1.0, one, z), M @ z, atol=1e-12)
    assert np.allclose(O.tea_apply(M, 0.0, one, z), np.diag(M) * z, atol=1e-12)
    Cv = np.random.default_rng(1).uniform(0.5, 1, 60)
    y = O.tea_apply(M, 0.3, Cv, z)
    D = np.diag(np.diag(M))
    assert np.allclose(y, Cv * (0.3 * (M - D) @ z + D @ z), atol=1e-12)


def test_jacobi_sqrt_squares_back():
    L = 16.0
    pos = random_walk(12, 2, L, 3)
    M = O.ewald_dense(O.Cell(L, 0, 0.1 * L, L, L), pos, 0.5)
    z = np.random.default_rng(2).normal(size=M.shape[0])
    y, st, lo, hi = O.sym_sqrt_apply(M, z)
    assert st == 0 and lo > 0
    yy, _, _, _ = O.sym_sqrt_apply(M, y)
    assert np.allclose(yy, M @ z, atol=1e-10)
    assert z @ M @ z == pytest.approx(y @ y, rel=1e-12)
    # numpy eigendecomposition as a second route
    w, V = np.linalg.eigh(M)
    assert np.allclose(y, V @ (np.sqrt(w) * (V.T @ z)), atol=1e-10)


def test_lanczos_oracle_converges_to_sqrt():
    L = 16.0
    pos = random_walk(20, 2, L, 3)
    M = O.ewald_dense(O.Cell(L, 0, 0.1 * L, L, L), pos, 0.5)
    z = np.random.default_rng(2).normal(size=M.shape[0])
    y, st, it = O.lanczos_sqrt(M, z, 1e-6, 60)
    ys, _, _, _ = O.sym_sqrt_apply(M, z)
    assert st == 0 and it < 40
    assert np.linalg.norm(y - ys) / np.linalg.norm(ys) < 1e-5


# ----------------------------------------------------------------- PME restatement
def test_spme_restatement_converges_to_dense_ewald():
    L = 16.0
    pos = random_walk(12, 2, L, 5)
    lat = O.lattice(1, L, 0.5, 1e-3, 0, 400)
    cell = (lat.L1xp, lat.L2xp, lat.L2yp, L)
    Me = O.ewald_dense(O.cell_of(lat), pos, 0.5)
    v = np.random.default_rng(4).normal(size=3 * len(pos))
    ur = Me @ v
    errs = []
    for K in ((16, 16, 16), (32, 32, 32), (64, 64, 64)):
        u = P.mobility_apply(pos, v, cell, 0.6, 7.0, K, 6)
        errs.append(np.linalg.norm(u - ur) / np.linalg.norm(ur))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 2e-5


def test_bspline_partition_of_unity():
    for p in (4, 6, 8):
        x = np.linspace(0, 0.999, 17)
        W = P.bspline_weights(x, p)
        assert np.allclose(W.sum(-1), 1.0)
        assert np.all(W >= 0)


# ----------------------------------------------------------------- MC
def mc_setup(nb, nc, L, seed, nq):
    pos = random_walk(nb, nc, L, seed)
    lat = O.lattice(1, L, 0.1, 1e-3, 1, 50)
    rx, ry, rz = O.apply_pbc(lat, *xyz(pos))
    rng = np.random.default_rng(seed)
    track = rng.choice(nb * nc, nq, replace=False).astype(np.int32)
    return lat, rx, ry, rz, track


def test_mc_two_charge_closed_form_and_draw_counts():
    # two beads on one chain of 3: bead0 charged; a right hop to bead1 is the only
    # option (mm > .5, no close non-adjacent bead), energy 0 with one charge
    lat = O.lattice(0, 40.0, 0.0, 1e-3)
    rx, ry, rz = np.array([0.0, 2.0, 4.0]), np.zeros(3), np.zeros(3)
    mc = O.ChargeHop(3, 1, [0])
    st, cur, log, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, 3.0, 3.0, np.array([0.7, 0.01], np.float32))
    # x = exp(-3) < 1 -> second draw 0.01 <= x accepts
    assert st == 0 and cur == 2 and list(log[0]) == [0, 0, 1, 0, 1, 2] and mc.track[0] == 1
    mc = O.ChargeHop(3, 1, [0])
    st, cur, log, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, 3.0, 3.0, np.array([0.2, 0.9], np.float32))
    # left of bead 0 is blocked: skipped with a single draw
    assert cur == 1 and log[0][3] == -1 and mc.track[0] == 0


def test_mc_pop_targets_lowest_close_bead():
    lat = O.lattice(0, 40.0, 0.0, 1e-3)
    # chain of 6 folded so beads 0 and 5 and 4 are within 2.1 of bead 2? build explicitly
    rx = np.array([0.0, 2.0, 4.0, 10.0, 5.0, 4.5])
    ry = np.array([0.0, 0.0, 0.0, 0.0, 1.5, 1.0])
    rz = np.zeros(6)
    mc = O.ChargeHop(6, 1, [2])
    st, cur, log, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, 0.0, 0.0, np.array([0.5], np.float32))
    # close non-adjacent beads of 2: bead 4 (r=1.80) and 5 (r=1.12); lowest index wins
    assert log[0][2] == 4 and log[0][3] == 2 and log[0][4] == 1 and mc.track[0] == 4


def test_mc_energy_bookkeeping_and_uniqueness():
    lat, rx, ry, rz, track = mc_setup(30, 4, 25.0, 7, 24)
    mc = O.ChargeHop(30, 4, track)
    u = O.Ran1(830062920).uniforms(200)
    st, cur, log, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, 1.0, 1.5, u)
    assert st == 0
    q = np.zeros(120, np.int32)
    q[track] = 1
    assert e0 == pytest.approx(O.coulomb_energy(lat, rx, ry, rz, q, 3.59, 5.0))
    assert len(set(mc.track.tolist())) == len(track)
    assert cur == sum(int(l[5]) for l in log)
    assert np.all(np.isin(log[:, 5], [1, 2]))


# ------------------------------------------------------------------ matrix-free Ewald
@pytest.mark.parametrize("flow", [0, 1, 2])
def test_ewald_matvec_equals_dense_product(flow):
    # the structure-factor arrangement equals the dense matrix for any split,
    # on deformed lattices and with overlapping beads
    L = 22.0
    lat = O.lattice(flow, L, 0.5 if flow else 0.0, 1e-3, 1, 300)
    pos = random_walk(30, 4, L, 3, min_sep=0.4)
    if flow == 2:
        rx, ry = O.rotate(lat, pos[:, 0].copy(), pos[:, 1].copy(), 0)
        pos = np.stack([rx, ry, pos[:, 2]], 1)
    cell = O.cell_of(lat)
    v = np.random.default_rng(1).normal(size=pos.size)
    ref = O.ewald_dense(cell, pos, 0.5) @ v
    for a in (None, 0.2, 0.9):
        u = O.ewald_matvec(cell, pos, v, alpha=a)
        assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-12


# ------------------------------------------------------------------ shear wrap shift
def test_shear_delx_shift_is_minus_gamma_above_and_plus_below():
    # applyPBC/getNID (S/main26.c:1409-1457): with the reference's compat shift, a
    # bead or pair vector whose y lies beyond +L2yp/2 moves by -delx = -edot*treal*dt
    # in x (modulo the cell), one beyond -L2yp/2 by +delx, others not at all
    L, edot, dt, t = 20.0, 0.4, 1e-3, 700
    ex = O.lattice(1, L, edot, dt, 0, t)
    cp = O.lattice(1, L, edot, dt, 1, t)
    delx = edot * (t % cp.tp) * dt
    assert delx > 0.1
    rng = np.random.default_rng(3)
    rx, ry, rz = (rng.uniform(-1.4 * L, 1.4 * L, 400) for _ in range(3))
    a = O.apply_pbc(cp, rx.copy(), ry.copy(), rz.copy())
    b = O.apply_pbc(ex, rx.copy(), ry.copy(), rz.copy())
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    d = a[0] - b[0]
    up, dn = ry > cp.L2yp / 2, ry < -cp.L2yp / 2
    k_up = (d[up] + delx) / cp.L1xp
    k_dn = (d[dn] - delx) / cp.L1xp
    assert np.allclose(k_up, np.round(k_up), atol=1e-9) and np.allclose(k_dn, np.round(k_dn), atol=1e-9)
    mid = ~(up | dn)
    assert np.array_equal(d[mid], np.zeros(mid.sum()))
    # pair vectors
    n = 60
    px, py, pz = (rng.uniform(-L / 2, L / 2, n) for _ in range(3))
    for i in range(0, n, 3):
        for j in range(1, n, 7):
            gc = O.nid(cp, px, py, pz, i, j)
            ge = O.nid(ex, px, py, pz, i, j)
            dy = py[i] - py[j]
            s = -1.0 if dy > cp.L2yp / 2 else (1.0 if dy < -cp.L2yp / 2 else 0.0)
            kk = (gc[0] - ge[0] - s * delx) / cp.L1xp
            assert abs(kk - round(kk)) < 1e-9 and gc[1] == ge[1] and gc[2] == ge[2]


# ------------------------------------------------------------------ bending force
def bend_energy(pos, nb, nc, kappab):
    """U = kappab theta^2 / 2 per interior bead, theta = pi - acos(cos) of the angle
    between r_j - r_{j-1} and r_j - r_{j+1} (S/main26.c:581-602)."""
    U = 0.0
    for c in range(nc):
        for j in range(1, nb - 1):
            b = c * nb + j
            a1, a2 = pos[b] - pos[b - 1], pos[b] - pos[b + 1]
            cs = a1 @ a2 / np.linalg.norm(a1) / np.linalg.norm(a2)
            U += 0.5 * kappab * (np.pi - np.arccos(cs)) ** 2
    return U


def test_bending_force_is_minus_gradient_of_bending_energy():
    L, nb, nc, kb = 200.0, 8, 2, 2.5
    lat = O.lattice(0, L, 0.0, 1e-3)
    rng = np.random.default_rng(5)
    steps = rng.normal(size=(nc, nb, 3))
    steps *= 2.0 / np.linalg.norm(steps, axis=2, keepdims=True)
    pos = (np.cumsum(steps, axis=1) + rng.uniform(-20, 20, (nc, 1, 3))).reshape(-1, 3)  # no images
    rx, ry, rz = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    st, f, _ = O.bond_force(lat, nb, nc, rx, ry, rz, 0, 0.0, 2.0, kb)  # Hookean with kappas = 0: bending only
    h = 1e-6
    g = np.zeros_like(pos)
    for b in range(nb * nc):
        for k in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[b, k] += h
            p2[b, k] -= h
            g[b, k] = (bend_energy(p1, nb, nc, kb) - bend_energy(p2, nb, nc, kb)) / (2 * h)
    assert np.abs(f.reshape(-1, 3) + g).max() < 1e-6 * max(1.0, np.abs(g).max())
    assert np.abs(g).max() > 0.1


# ------------------------------------------------------------------ chargeHop branches
def hop_geometry():
    """One chain of 6 in a large box; bead 2 has two close non-adjacent beads (r < 2.1):
    bead 4 at r = 2.06 and bead 5 at r = 1.68; no other pair of the chain is that close
    except along bonds."""
    rx = np.array([0.0, 2.0, 4.0, 6.0, 5.0, 3.5])
    ry = np.array([0.0, 0.0, 0.0, 0.0, 1.8, 1.6])
    rz = np.zeros(6)
    return O.lattice(0, 40.0, 0.0, 1e-3), rx, ry, rz


HOP_CASES = [
    # (charges, uniforms, barrier, barrier2, expected log[0] = {k, from, to, kind, accepted, draws})
    # mm < .33 with a close bead: right neighbour, "adjacent near" kind, exp(0) = 1 accepts on one draw
    ([2], [0.2], 0.0, 5.0, [0, 2, 3, 1, 1, 1]),
    # mm > .66 with a close bead: left neighbour; exp(-0.5) < 1 so a dice is drawn: 0.9 rejects, 0.3 accepts
    ([2], [0.8, 0.9], 0.5, 5.0, [0, 2, 1, 1, 0, 2]),
    ([2], [0.8, 0.3], 0.5, 5.0, [0, 2, 1, 1, 1, 2]),
    # .33 <= mm <= .66: pop to the lowest-index close bead (4), with barrier2 (0) not barrier (5)
    ([2], [0.5], 5.0, 0.0, [0, 2, 4, 2, 1, 1]),
    # pop blocked (bead 4 charged): the flag stays set, so bead 5 and the neighbour fallback are
    # blocked too and the charge is skipped after one draw
    ([2, 4], [0.5, 0.9, 0.9, 0.9], 0.0, 0.0, [0, 2, -1, -1, 0, 1]),
    # mm < .33 with the right neighbour charged: blocked, and the fallback to the left is blocked too
    ([2, 3], [0.1, 0.9, 0.9, 0.9], 0.0, 0.0, [0, 2, -1, -1, 0, 1]),
    # no close bead for bead 0 (chain start): mm > .5 hops right without the near kind
    ([0], [0.7], 0.0, 0.0, [0, 0, 1, 0, 1, 1]),
    # chain end (bead 5) cannot hop right: skipped
    ([5], [0.2], 0.0, 0.0, None),
]


@pytest.mark.parametrize("case", range(len(HOP_CASES)))
def test_chargehop_branches_hand_worked(case):
    track, u, b1, b2, expect = HOP_CASES[case]
    lat, rx, ry, rz = hop_geometry()
    mc = O.ChargeHop(6, 1, track)
    st, cur, log, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, b1, b2, np.array(u, np.float32))
    if expect is None:
        # bead 5: close to bead 2 (r = 1.68), mm = 0.2 < .33 asks for the right neighbour,
        # which does not exist on a chain end: skipped, fallback left also blocked by the flag
        assert log[0][3] == -1 and log[0][4] == 0 and mc.track[0] == 5
        return
    got = list(log[0])
    if expect[2] == -1:  # skipped: the target column is not meaningful
        got[2] = -1
    assert got == expect
    moved = expect[4] == 1
    assert mc.track[0] == (expect[2] if moved else track[0])
    if expect[2] >= 0 and not moved:
        assert mc.old[0] == track[0]


def test_bending_stress_closed_forms():
    """The bending terms of the stress accumulators (S/main26.c:606-613) at two
    geometries worked by hand.  Bead j has getNID vectors u = r_j - r_{j-1} and
    w = r_j - r_{j+1}; cos = u.w/(|u||w|), theta = pi - acos(cos) and
    coeff = -kappa_b theta / sqrt(1 - cos^2).
    (1) u = (a,0,0), w = (0,b,0): cos = 0, theta = pi/2; of the four tauxx terms
        only -u_x coeff u_x/(|u||w|) survives: tauxx = kappa_b pi a / (2 b); every
        tauyy term carries u_y = 0 or w_z = 0 or cos = 0: tauyy = 0.
    (2) u = (0,1,2), w = (0,0,1): |u| = sqrt5, |w| = 1, cos = 2/sqrt5; the tauyy
        terms give coeff [2/(5 sqrt5) - 2 (1/sqrt5 - 2/(5 sqrt5))] = -4 coeff/(5 sqrt5),
        so tauyy = 4 kappa_b theta / 5; its second term reads u_z, where a y-y
        version would give coeff (2 - 3)/(5 sqrt5) instead.
    The springs are Hookean with kappas = 0, so only bending contributes."""
    L, kb = 40.0, 1.7
    lat = O.lattice(0, L, 0.0, 1e-3)
    c = np.array([1.3, -0.7, 0.4])
    a, b = 1.5, 2.5
    for (u, w), expect in (
            (((a, 0.0, 0.0), (0.0, b, 0.0)), (kb * np.pi * a / (2 * b), 0.0, 0.0)),
            (((0.0, 1.0, 2.0), (0.0, 0.0, 1.0)), (0.0, 4 * kb * (np.pi - np.arccos(2 / np.sqrt(5))) / 5, 0.0))):
        p = np.array([c - np.array(u), c, c - np.array(w)])
        rx, ry, rz = (np.ascontiguousarray(p[:, k]) for k in range(3))
        st, f, tau = O.bond_force(lat, 3, 1, rx, ry, rz, 0, 0.0, 2.0, kb)
        assert st == 0
        assert np.allclose(tau, expect, rtol=1e-12, atol=1e-12), (tau, expect)
