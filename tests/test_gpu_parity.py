"""GPU parity tests: the HIP library through its C ABI against the CPU oracle.
Bit-exact for integer and index results (remap, min image, bins, lists, MC
sequence); tolerances stated per test for floating point."""
import numpy as np
import pytest
from helpers import O, S, random_walk

import spme_ref as P

pytestmark = pytest.mark.gpu


def xyz(pos):
    return [np.ascontiguousarray(pos[:, k]) for k in range(3)]


def ctx_for(flow, L, nb, nc, prec=S.PREC_MIXED, compat=1, edot=0.3, dt=1e-3, **kw):
    p = S.default_params(nb=nb, nc=nc, L=L, flow_type=flow, edot=edot, dt=dt, precision=prec,
                         pbc_compat=compat, **kw)
    return S.Context(p), p


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_remap(lat, t, rx, ry, rz):
    """The reference's per-step remap sequence (X/main14.c:69-78, S/main26.c:71-73)."""
    if lat.flow == 2:
        rx, ry = O.rotate(lat, rx, ry, 0)
        rx, ry = O.rotate(lat, rx, ry, 1)
    O.lattice_update(lat, t)
    if lat.flow == 2:
        rx, ry = O.rotate(lat, rx, ry, 0)
    return O.apply_pbc(lat, rx, ry, rz)


# ----------------------------------------------------------------- remap / images / lists
@pytest.mark.parametrize("flow,compat", [(1, 1), (1, 0), (2, 1), (0, 1)])
def test_remap_bit_exact(flow, compat):
    L = 26.0
    ctx, p = ctx_for(flow, L, 30, 5, compat=compat, edot=0.4)
    lat = O.lattice(flow, L, 0.4 if flow else 0.0, 1e-3, compat)
    pos = random_walk(30, 5, L, 3) * 1.7  # some beads well outside the box
    rx, ry, rz = xyz(pos)
    ctx.upload_positions(rx, ry, rz)
    rng = np.random.default_rng(0)
    for t in (0, 1, 2, 700, 701, lat.tp - 1 if lat.tp < 10**8 else 3, lat.tp if lat.tp < 10**8 else 4):
        ctx.remap(t)
        rx, ry, rz = oracle_remap(lat, t, rx, ry, rz)
        gx, gy, gz = ctx.download_positions()
        assert np.array_equal(gx, rx) and np.array_equal(gy, ry) and np.array_equal(gz, rz), t
        # move beads (as a step would) and continue from identical state
        d = rng.normal(scale=0.7, size=(len(rx), 3))
        rx, ry, rz = rx + d[:, 0], ry + d[:, 1], rz + d[:, 2]
        ctx.upload_positions(rx, ry, rz)
        if flow == 2:  # positions return to the lab frame between steps
            rx, ry = O.rotate(lat, rx, ry, 1)
            ctx.rotate(1)
            gx, gy, _ = ctx.download_positions()
            assert np.array_equal(gx, rx) and np.array_equal(gy, ry)


@pytest.mark.parametrize("flow,compat", [(1, 1), (2, 1), (1, 0)])
def test_min_image_and_bins_bit_exact(flow, compat):
    L = 30.0
    ctx, p = ctx_for(flow, L, 40, 5, compat=compat, edot=0.2)
    lat = O.lattice(flow, L, 0.2, 1e-3, compat)
    binsz = O.bins(lat, 2.0)
    rx, ry, rz = xyz(random_walk(40, 5, L, 4))
    ctx.upload_positions(rx, ry, rz)
    t = lat.tp // 2 + 37  # second half of a period: nmax = 2 in getNIDbin
    for tt in (5, t):
        ctx.remap(tt)
        rx, ry, rz = oracle_remap(lat, tt, rx, ry, rz)
        rng = np.random.default_rng(tt)
        i = rng.integers(0, 200, 500).astype(np.int32)
        j = rng.integers(0, 200, 500).astype(np.int32)
        g = ctx.min_image(0, i, j)
        o = np.array([O.nid(lat, rx, ry, rz, a, b) for a, b in zip(i, j)])
        assert np.array_equal(g, o)
        gb = ctx.bin_index(0, 2.0, i, j)
        ob = np.array([O.binij(binsz, lat, rx, ry, rz, a, b) for a, b in zip(i, j)])
        assert np.array_equal(gb, ob)


@pytest.mark.parametrize("flow,compat,nb,nc,L", [(1, 1, 40, 6, 24.0), (1, 0, 50, 4, 20.0), (2, 1, 40, 6, 24.0),
                                                 (1, 1, 100, 1, 60.0)])
def test_verlet_list_bit_exact(flow, compat, nb, nc, L):
    ctx, p = ctx_for(flow, L, nb, nc, compat=compat, edot=0.3)
    lat = O.lattice(flow, L, 0.3, 1e-3, compat)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 5))
    ctx.upload_positions(rx, ry, rz)
    for t in (0, 333):
        ctx.remap(t)
        rx, ry, rz = oracle_remap(lat, t, rx, ry, rz)
        ctx.neighbors(True)
        nl, lst = ctx.verlet(0)
        onl, olst = O.verlet(lat, rx, ry, rz, 6.0)
        assert np.array_equal(nl, onl)
        for a in range(len(nl)):
            assert np.array_equal(lst[a, :nl[a]], olst[a, :onl[a]])


def test_verlet_trigger_follows_reference():
    """checkVerlet (S/main26.c:858-869): the lists are rebuilt when some bead's
    displacement summed as a vector since the last build (dxv, S/main26.c:726,742,849)
    exceeds rnew = (rv - rc)/2 = 0.5 (S/Initialization.c:303-307).  Force-free
    free-draining steps move bead 0 by exactly sqrt(2 dt) z."""
    L, dt = 30.0, 1e-3
    ctx, p = ctx_for(1, L, 2, 8, edot=0.0, dt=dt, kappas=0.0, epsilon=0.0, noise_mode=S.NOISE_FREE_DRAINING,
                     prec=S.PREC_FP64)
    pos = np.array([[x, y, z] for x in (-10.5, -3.5, 3.5, 10.5) for y in (-7.5, 7.5) for z in (-7.5, 7.5)])
    ctx.upload_positions(*xyz(pos))
    assert ctx.neighbors(False) == 1  # first call builds
    assert ctx.neighbors(False) == 0
    n = len(pos)

    def move(dx):  # one step moving bead 0 by dx in x, every other bead by 0
        z = np.zeros(3 * n)
        z[0] = dx / np.sqrt(2 * dt)
        ctx.step(1, z)
        return ctx.neighbors(False)

    # vector sum: +0.3, -0.3, +0.45 travels 1.05 but sits 0.45 from the build
    assert [move(0.3), move(-0.3), move(0.45)] == [0, 0, 0]
    assert move(0.04) == 0  # 0.49
    assert move(0.02) == 1  # 0.51 > 0.5: rebuilt, sums reset
    assert move(0.49) == 0
    assert move(-0.98) == 0  # |-0.49|
    assert move(-0.02) == 1
    gx, _, _ = ctx.download_positions()
    assert abs(gx[0] - (pos[0, 0] + 0.3 - 0.3 + 0.45 + 0.04 + 0.02 + 0.49 - 0.98 - 0.02)) < 1e-12
    ctx.close()


# ----------------------------------------------------------------- forces
@pytest.mark.parametrize("flow,fene", [(1, 0), (2, 0), (1, 1)])
def test_forces_match_oracle(flow, fene):
    L = 28.0
    nb, nc = 40, 5
    kw = dict(spring_fene=fene, kappas=30.0 if fene else 200.0, qmax=3.0 if fene else 2.0, kappab=0.7,
              lj_ext_form=1 if flow == 2 else 0)
    ctx, p = ctx_for(flow, L, nb, nc, edot=0.2, **kw)
    lat = O.lattice(flow, L, 0.2, 1e-3, 1)
    pos = random_walk(nb, nc, L, 7, bond=2.0 if not fene else 2.4)
    rx, ry, rz = xyz(pos)
    ctx.upload_positions(rx, ry, rz)
    rng = np.random.default_rng(1)
    track = rng.choice(nb * nc, 60, replace=False).astype(np.int32)
    sign = np.where(np.arange(60) % 2 == 0, 1, -1).astype(np.int32)
    ctx.upload_charges([60], track, sign)
    ctx.remap(50)
    rx, ry, rz = oracle_remap(lat, 50, rx, ry, rz)
    ctx.neighbors(True)
    f = ctx.forces()
    st, fb, _ = O.bond_force(lat, nb, nc, rx, ry, rz, fene, kw["kappas"], kw["qmax"], kw["kappab"])
    assert st == 0
    nl, lst = O.verlet(lat, rx, ry, rz, 6.0)
    fl, _ = O.lj_force(lat, rx, ry, rz, nl, lst, 5.03968 if flow == 2 else 25.0, 0.1, flow == 2)
    q = np.zeros(nb * nc, np.int32)
    q[track] = sign
    fc = O.coulomb_force(lat, rx, ry, rz, q, 3.59, 5.0)
    ref = fb + fl + fc
    assert np.abs(f - ref).max() <= 1e-10 * np.abs(ref).max()


def test_fene_overstretch_is_reported():
    L = 30.0
    ctx, p = ctx_for(1, L, 3, 1, spring_fene=1, kappas=30.0, qmax=2.2)
    ctx.upload_positions(np.array([0.0, 2.0, 6.0]), np.zeros(3), np.zeros(3))
    with pytest.raises(S.SmdltError) as e:
        ctx.forces()
    assert e.value.code == -3


# ----------------------------------------------------------------- mobility
def mob_case(flow=1, L=24.0, nb=40, nc=3, seed=1, t=300, prec=S.PREC_MIXED, **kw):
    ctx, p = ctx_for(flow, L, nb, nc, prec=prec, edot=0.5, **kw)
    lat = O.lattice(flow, L, 0.5, 1e-3, 1, t)
    pos = random_walk(nb, nc, L, seed)
    rx, ry, rz = xyz(pos)
    if flow == 2:  # the mobility works in the wrap frame
        rx, ry = O.rotate(lat, rx, ry, 0)
    ctx.upload_positions(rx, ry, rz)
    ctx.update_lattice(t)
    return ctx, lat, np.stack([rx, ry, rz], 1)


@pytest.mark.parametrize("flow", [1, 2])
def test_mobility_fp64_equals_pme_restatement(flow):
    ctx, lat, pos = mob_case(flow, prec=S.PREC_FP64)
    cfg = ctx.mobility_config()
    v = np.random.default_rng(2).normal(size=pos.size)
    u = ctx.mobility_apply(v)
    un = P.mobility_apply(pos, v, (lat.L1xp, lat.L2xp, lat.L2yp, lat.L), cfg["alpha"], cfg["rc"], cfg["grid"],
                          cfg["order"])
    assert rel(u, un) < 1e-12


@pytest.mark.parametrize("grid", [(40, 48, 72), (24, 80, 120), (96, 32, 40), (160, 144, 16), (16, 40, 160),
                                  (16, 16, 384), (16, 24, 320)])
def test_mobility_mixed_radix_grids_equal_pme_restatement(grid):
    # line lengths 2^a 3^b 5^c exercise the radix-3 and radix-5 Stockham stages; planes
    # with K0*K1/2 <= 8192 take the fused x-y pass, (160, 144, 16) the separate x and y passes
    # (as do the four 128^3 grids of the large-grid batch test: over 8e6 points in total)
    ctx, lat, pos = mob_case(1, prec=S.PREC_FP64, grid=grid, spline_order=6)
    cfg = ctx.mobility_config()
    assert cfg["grid"] == grid
    v = np.random.default_rng(12).normal(size=pos.size)
    u = ctx.mobility_apply(v)
    un = P.mobility_apply(pos, v, (lat.L1xp, lat.L2xp, lat.L2yp, lat.L), cfg["alpha"], cfg["rc"], cfg["grid"],
                          cfg["order"])
    assert rel(u, un) < 1e-12


@pytest.mark.parametrize("flow,prec", [(1, S.PREC_MIXED), (1, S.PREC_FP64), (2, S.PREC_MIXED), (0, S.PREC_MIXED)])
def test_mobility_matches_dense_ewald(flow, prec):
    # M.F against the exact dense FP64 Ewald sum: <= 1e-5 relative
    ctx, lat, pos = mob_case(flow, prec=prec)
    Md = O.ewald_dense(O.cell_of(lat), pos, 0.5)
    for seed in (2, 3):
        v = np.random.default_rng(seed).normal(size=pos.size)
        assert rel(ctx.mobility_apply(v), Md @ v) < 1e-5


@pytest.mark.parametrize("grid", [(16, 16, 384), (16, 24, 320)])
def test_mobility_long_z_lines_few_tasks_mixed(grid):
    # few z lines of a length above the smallest z tile (256 values): the z pass
    # must keep a tile that holds a whole line
    ctx, lat, pos = mob_case(1, grid=grid, spline_order=6)
    cfg = ctx.mobility_config()
    assert cfg["grid"] == grid
    v = np.random.default_rng(14).normal(size=pos.size)
    un = P.mobility_apply(pos, v, (lat.L1xp, lat.L2xp, lat.L2yp, lat.L), cfg["alpha"], cfg["rc"], cfg["grid"],
                          cfg["order"])
    assert rel(ctx.mobility_apply(v), un) < 1e-5


def test_mobility_split_invariance():
    vals = []
    v = np.random.default_rng(5).normal(size=3 * 120)
    for rc in (6.0, 9.0):
        ctx, lat, pos = mob_case(1, rc_real=rc)
        vals.append(ctx.mobility_apply(v))
    assert rel(vals[0], vals[1]) < 2e-5


@pytest.mark.parametrize("rc", [80.0, 85.0])
def test_mobility_long_cutoff_tables_match_fp64(rc):
    # the cluster pass stages coefficient tables of up to 1280 entries (rc <= 81) in LDS
    # (rc = 80: the largest such table) and reads longer ones from global memory (rc = 85);
    # the FP64 path at the same split evaluates the closed forms
    mixed, lat, pos = mob_case(0, L=180.0, rc_real=rc)
    fp64, _, _ = mob_case(0, L=180.0, rc_real=rc, prec=S.PREC_FP64)
    # mixed-precision cut-offs are raised to the next table knot (r^2 on a segment boundary)
    assert rc <= mixed.mobility_config()["rc"] < rc * (1 + 1 / 64)
    assert fp64.mobility_config()["rc"] == rc
    v = np.random.default_rng(13).normal(size=pos.size)
    assert rel(mixed.mobility_apply(v), fp64.mobility_apply(v)) < 1e-5


def _images(cell, n=2):
    r = np.arange(-n, n + 1)
    i, j, k = [a.ravel() for a in np.meshgrid(r, r, r, indexing="ij")]
    return np.stack([i * cell.a1x + j * cell.a2x, i * cell.a1y + j * cell.a2y, k * cell.Lz], 1)


def _half_width(cell):
    """Half the shortest lattice vector bound the cluster pass uses: half the smaller
    of the reduced 2D cell's smallest width and Lz."""
    b1, b2 = np.array([cell.a1x, cell.a1y]), np.array([cell.a2x, cell.a2y])
    while True:  # Lagrange reduction
        if b1 @ b1 > b2 @ b2:
            b1, b2 = b2, b1
        m = np.rint((b1 @ b2) / (b1 @ b1))
        if m == 0:
            break
        b2 = b2 - m * b1
    area = abs(b1[0] * b2[1] - b1[1] * b2[0])
    return 0.5 * min(area / np.linalg.norm(b2), cell.Lz)


@pytest.mark.parametrize("flow,rc", [(1, 6.0), (1, 8.0), (2, 6.0), (2, 8.0)])
def test_cluster_image_shift_paths_match_dense_ewald(flow, rc):
    # The cluster pass sums a cluster pair under one image shift when the first
    # beads' minimum-image distance plus both cluster radii is at most 2h - rc
    # (h: half the shortest lattice vector), and takes each pair's own minimum
    # image otherwise. This box holds cluster pairs, with a bead pair within rc,
    # on both sides of that bound; M.F against the dense FP64 Ewald sum.
    nb, nc, L = 40, 3, 24.0
    ctx, lat, pos = mob_case(flow, L=L, nb=nb, nc=nc, rc_real=rc)
    cell = O.cell_of(lat)
    rce = ctx.mobility_config()["rc"]
    img = _images(cell)
    first = [(ch * nb + j0, min(8, nb - j0)) for ch in range(nc) for j0 in range(0, nb, 8)]

    def mi(d):  # minimum-image lengths of vectors d (..., 3)
        return np.linalg.norm(d[..., None, :] + img, axis=-1).min(-1)

    # cluster radius: the largest minimum-image distance from the cluster's first bead
    rad = [mi(pos[f:f + k] - pos[f]).max() for f, k in first]
    smax = 2 * _half_width(cell) - rce
    n_one = n_own = 0
    for I, (fi, ki) in enumerate(first):
        for J, (fj, kj) in enumerate(first):
            if mi(pos[fj:fj + kj, None, :] - pos[None, fi:fi + ki, :]).min() >= rce:
                continue  # no bead pair within the cut-off
            if mi(pos[fj] - pos[fi]) + rad[I] + rad[J] < smax:
                n_one += 1
            else:
                n_own += 1
    assert n_one > 0 and n_own > 0, (n_one, n_own)
    Md = O.ewald_dense(cell, pos, 0.5)
    for seed in (4, 5):
        v = np.random.default_rng(seed).normal(size=pos.size)
        assert rel(ctx.mobility_apply(v), Md @ v) < 1e-5


@pytest.mark.parametrize("order,grid", [(4, (48, 40, 32)), (6, (48, 40, 32)), (8, (48, 40, 32)),
                                        (8, (16, 24, 40)), (6, (40, 16, 16))])
@pytest.mark.parametrize("flow", [1, 2])
def test_matrix_core_spread_matches_fp64_gather(flow, order, grid):
    # the mixed path spreads on the matrix cores in FP32, the FP64 path gathers per
    # grid point (and equals the PME restatement to 1e-12); at the same split, grid
    # and spline order the two products agree to the mixed path's FP32 precision
    # (grids with an axis of two bricks make each brick's +1 neighbour wrap around)
    kw = dict(grid=grid, spline_order=order, rc_real=8.0)
    mixed, lat, pos = mob_case(flow, nb=60, nc=4, **kw)
    fp64, _, _ = mob_case(flow, nb=60, nc=4, prec=S.PREC_FP64, **kw)
    cm, cf = mixed.mobility_config(), fp64.mobility_config()
    assert cm["order"] == cf["order"] == order and list(cm["grid"]) == list(cf["grid"]) == list(grid)
    assert cm["alpha"] == cf["alpha"]
    v = np.random.default_rng(21).normal(size=pos.size)
    e = rel(mixed.mobility_apply(v), fp64.mobility_apply(v))
    print(f"mixed (matrix-core spread) vs FP64 product, flow {flow}, p = {order}, grid {grid}: {e:.2e}")
    assert e < 1e-5


def test_mobility_symmetric():
    ctx, lat, pos = mob_case(1)
    rng = np.random.default_rng(6)
    v, w = rng.normal(size=pos.size), rng.normal(size=pos.size)
    a, b = w @ ctx.mobility_apply(v), v @ ctx.mobility_apply(w)
    assert abs(a - b) <= 1e-6 * np.linalg.norm(v) * np.linalg.norm(w)


def test_single_bead_hasimoto():
    L = 20.0
    ctx, p = ctx_for(0, L, 1, 1, prec=S.PREC_FP64)
    ctx.upload_positions(np.array([0.3]), np.array([-1.2]), np.array([2.7]))
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1
        u = ctx.mobility_apply(e)
        assert u[k] == pytest.approx(1 - 2.8372974795 / L + 4.1887902048 / L ** 3, rel=1e-5)


# ----------------------------------------------------------------- Brownian
@pytest.mark.parametrize("prec,mtol", [(S.PREC_MIXED, 1e-5), (S.PREC_FP64, 1e-5), (S.PREC_FP64, 1e-7)])
def test_lanczos_matches_dense_sqrt(prec, mtol):
    # M^1/2 z within 1e-5 of the dense square root (Jacobi) at Lanczos tol 1e-6;
    # FP64 also with the mobility settings tuned to 1e-7
    ctx, lat, pos = mob_case(1, prec=prec, mobility_tol=mtol)
    Md = O.ewald_dense(O.cell_of(lat), pos, 0.5)
    z = np.random.default_rng(3).normal(size=pos.size)
    y, it, res, code = ctx.brownian(z, 1e-6, 30)
    assert code == 0 and 2 <= it[0] <= 30
    ys, st, lo, hi = O.sym_sqrt_apply(Md, z)
    e = rel(y, ys)
    print(f"M^1/2 z vs dense square root, precision {prec}, mobility_tol {mtol}: {e:.2e}")
    assert e < 1e-5
    assert y @ y == pytest.approx(z @ Md @ z, rel=1e-5)


def test_lanczos_long_solve_matches_dense_sqrt():
    # an unreachable tolerance runs the solve to 100 iterations (tridiagonal
    # square roots of order > 64); the result still equals the dense square root
    # (300 overlapping-walk beads in L = 20: condition number ~130)
    ctx, lat, pos = mob_case(1, L=20.0, nb=60, nc=5, prec=S.PREC_FP64, lanczos_maxit=100)
    Md = O.ewald_dense(O.cell_of(lat), pos, 0.5)
    z = np.random.default_rng(13).normal(size=pos.size)
    y, it, res, code = ctx.brownian(z, 1e-15, 100)
    assert code == -4 and it[0] == 100
    w, Q = np.linalg.eigh(Md)
    ys = Q @ (np.sqrt(w) * (Q.T @ z))
    assert rel(y, ys) < 2e-5  # the PME product itself is accurate to ~1e-5 (mobility_tol)


def test_lanczos_zero_noise_gives_zero():
    ctx, lat, pos = mob_case(1)
    y, it, res, code = ctx.brownian(np.zeros(pos.size), 1e-6, 30)
    assert code == 0 and np.all(y == 0.0)


def test_tea_matches_oracle():
    ctx, lat, pos = mob_case(1)
    Md = O.ewald_dense(O.cell_of(lat), pos, 0.5)
    rng = np.random.default_rng(4)
    z = rng.normal(size=pos.size)
    Cv = rng.uniform(0.6, 1.0, pos.size)
    y = ctx.tea_apply(0.37, Cv, z)
    assert rel(y, O.tea_apply(Md, 0.37, Cv, z)) < 1e-5


@pytest.mark.parametrize("flow,t", [(1, 100), (1, 15000), (0, 0), (2, 300)])
def test_tea_sample_matches_ewald_bin_catea(flow, t):
    """ewaldBin + CATEA (S/main26.c:975-1381): beta, C and eps of one sample against
    the oracle, in both halves of the shear period, at equilibrium and under KRBC;
    the random walks overlap, so the reference's r <= 2 branch is exercised."""
    L, nb, nc = 20.0, 20, 4
    n = nb * nc
    edot = 0.05 if flow else 0.0
    ctx, p = ctx_for(flow, L, nb, nc, prec=S.PREC_FP64, edot=edot, dt=1e-3)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 21))
    ctx.upload_positions(rx, ry, rz)
    ctx.update_lattice(t)
    beta, Cv, eps = ctx.tea_sample()
    lat = O.lattice(flow, L, edot, 1e-3, 1, t)
    e, rs, dtr, kc = O.ewald_bin(lat, rx, ry, rz, nb, L ** 3)
    b, Co = O.catea(n, e, rs)
    assert beta[0] == pytest.approx(b, rel=1e-10)
    assert eps[0] == pytest.approx(e / (9 * n * n - 3 * n), rel=1e-10)
    assert np.abs(Cv - Co).max() < 1e-10


def test_tea_sample_batched_equals_single_contexts():
    specs = [(1, 22.0, 20, 3), (2, 18.0, 15, 4), (0, 16.0, 10, 2)]
    params, poss = [], []
    for k, (flow, L, nb, nc) in enumerate(specs):
        params.append(S.default_params(nb=nb, nc=nc, L=L, flow_type=flow, edot=0.05 if flow else 0.0, dt=1e-3,
                                       precision=S.PREC_FP64))
        poss.append(random_walk(nb, nc, L, 50 + k))
    big = S.Context(params)
    allp = np.concatenate(poss)
    big.upload_positions(*xyz(allp))
    big.update_lattice(400)
    bb, Cb, eb = big.tea_sample()
    off = 0
    for k, pr in enumerate(params):
        one = S.Context(pr)
        one.upload_positions(*xyz(poss[k]))
        one.update_lattice(400)
        b1, C1, e1 = one.tea_sample()
        m = 3 * len(poss[k])
        assert bb[k] == pytest.approx(b1[0], rel=1e-12) and eb[k] == pytest.approx(e1[0], rel=1e-12)
        assert np.abs(Cb[off:off + m] - C1).max() < 1e-12
        off += m
        one.close()
    big.close()


@pytest.mark.parametrize("flow", [1, 2])
def test_far_table_matches_gridhi(flow, tmp_path):
    """gridHI (S/gridHI.c:3-278): every phase and bin against the restatement, and
    the mm/G text written by smdlt_write_far_table reads back in the reference's
    layout (S/Initialization.c:497-520)."""
    L, nt_avg = 14.0, 4
    ctx, p = ctx_for(flow, L, 10, 2, prec=S.PREC_FP64, edot=0.05, dt=1e-3)
    tab, dims, sizes = ctx.far_table(0, nt_avg, 2.0)
    lat = O.lattice(flow, L, 0.05, 1e-3, 1, 0)
    b = O.bins(lat, 2.0)
    assert dims == (b.nbx, b.nby, b.nbz)
    assert np.allclose(sizes, [b.bin_size_x, b.bin_size_y, b.bin_size_z], rtol=0, atol=0)
    ref = O.far_table(lat, nt_avg, b, L ** 3)
    assert np.abs(tab - ref).max() < 1e-11 * np.abs(ref).max()
    path = tmp_path / "G.txt"
    S.write_far_table(path, tab, dims)
    lines = path.read_text().splitlines()
    nb = int(np.prod(dims))
    assert len(lines) == 4 * nt_avg * nb
    k = 4 * (nt_avg - 1) * nb + 4 * 5  # phase nt_avg-1, bin 5
    i, j, kz = 5 // (dims[1] * dims[2]), (5 // dims[2]) % dims[1], 5 % dims[2]
    assert lines[k] == f"{i} {j} {kz} 1"
    back = np.array([[float(x) for x in lines[k + 1 + a].split()] for a in range(3)])
    assert np.allclose(back, tab[nt_avg - 1, 5], rtol=1e-6, atol=1e-12)


# ----------------------------------------------------------------- timestep
@pytest.mark.parametrize("flow,mode", [(1, S.NOISE_LANCZOS), (2, S.NOISE_LANCZOS), (1, S.NOISE_FREE_DRAINING),
                                       (2, S.NOISE_FREE_DRAINING), (1, S.NOISE_TEA), (2, S.NOISE_TEA)])
def test_step_matches_dense_update(flow, mode):
    """One step of updateHI (S/main26.c:745-857) with Lanczos or TEA noise, or of updateFD
    (S/main26.c:731-744), against the dense-matrix update."""
    L, nb, nc, dt, edot = 24.0, 30, 3, 1e-3, 0.5
    ctx, p = ctx_for(flow, L, nb, nc, prec=S.PREC_FP64, edot=edot, dt=dt, kappas=200.0, noise_mode=mode,
                     mobility_tol=1e-7)
    lat = O.lattice(flow, L, edot, dt, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 9))
    ctx.upload_positions(rx, ry, rz)
    rng = np.random.default_rng(8)
    z = rng.normal(size=3 * nb * nc)
    beta, Cv = 0.37, rng.uniform(0.6, 1.0, 3 * nb * nc)
    if mode == S.NOISE_TEA:
        ctx.set_tea(beta, Cv)
    t = 10
    ctx.step(t, z)
    gx, gy, gz = ctx.download_positions()
    # oracle: remap, forces, dense M, M F, M^(1/2) z (Jacobi), flow, back to lab frame
    wx, wy, wz = oracle_remap(lat, t, rx, ry, rz)
    nl, lst = O.verlet(lat, wx, wy, wz, 6.0)
    st, fb, _ = O.bond_force(lat, nb, nc, wx, wy, wz, 0, 200.0, 2.0)
    fl, _ = O.lj_force(lat, wx, wy, wz, nl, lst, 25.0, 0.1, 0)
    F = fb + fl  # lab frame
    pos = np.stack([wx, wy, wz], 1)
    if mode == S.NOISE_FREE_DRAINING:
        M = np.eye(pos.size)
    else:
        M = O.ewald_dense(O.cell_of(lat), pos, 0.5)
    if flow == 2:
        c, s = lat.ct, lat.st
        Rot = np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])  # lab -> wrap
        Fw = F @ Rot.T
        zw = (z.reshape(-1, 3) @ Rot.T).reshape(-1)
    else:
        Fw, zw = F, z
    if mode == S.NOISE_TEA:  # C and beta act on wrap-frame components
        ys = O.tea_apply(M, beta, Cv, zw)
    else:
        ys, _, _, _ = O.sym_sqrt_apply(M, zw)
    d = (dt * (M @ Fw.reshape(-1)) + np.sqrt(2 * dt) * ys).reshape(-1, 3)
    if flow == 2:
        RotI = Rot.T
        lab = pos @ RotI.T
        d = d @ RotI.T
        ex = lab[:, 0] + d[:, 0] + dt * edot * lab[:, 0]
        ey = lab[:, 1] + d[:, 1] - dt * edot * lab[:, 1]
        ez = lab[:, 2] + d[:, 2]
    else:
        ex = wx + d[:, 0] + dt * edot * wy
        ey = wy + d[:, 1]
        ez = wz + d[:, 2]
    disp_ref = np.stack([ex - (lab[:, 0] if flow == 2 else wx), ey - (lab[:, 1] if flow == 2 else wy),
                         ez - wz], 1)
    disp_gpu = np.stack([gx - (lab[:, 0] if flow == 2 else wx), gy - (lab[:, 1] if flow == 2 else wy),
                         gz - wz], 1)
    e = rel(disp_gpu, disp_ref)
    print(f"FP64 step vs dense update, flow {flow}, noise mode {mode}: {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("prec", [S.PREC_FP64, S.PREC_MIXED])
def test_graph_replay_matches_eager(prec):
    # replayed steps run the same kernels on the same data in the same order as
    # eager ones (a converged replica's further Lanczos iterations are skipped
    # by every kernel), so the trajectories agree bit for bit
    L, nb, nc = 30.0, 50, 4
    out = []
    for graph in (False, True):
        ctx, p = ctx_for(1, L, nb, nc, prec=prec, edot=0.3, dt=5e-4)
        rx, ry, rz = xyz(random_walk(nb, nc, L, 10))
        ctx.upload_positions(rx, ry, rz)
        ctx.run(0, 6, use_graph=graph)
        out.append(np.stack(ctx.download_positions(), 1))
        ctx.close()
    assert np.array_equal(out[0], out[1])


def test_nonfinite_displacement_is_reported_and_not_applied():
    # Hookean springs so stiff that the forces overflow: the free-draining step
    # reports SMDLT_E_NONFINITE and leaves the affected beads where they were
    L, nb, nc = 30.0, 10, 2
    ctx, p = ctx_for(0, L, nb, nc, prec=S.PREC_FP64, edot=0.0, dt=1e-3, kappas=1.7e308, qmax=0.0,
                     noise_mode=S.NOISE_FREE_DRAINING)  # |F| = kappas |d|: bond components above 1.06 overflow
    pos = random_walk(nb, nc, L, 3)
    ctx.upload_positions(*xyz(pos))
    with pytest.raises(S.SmdltError) as e:
        ctx.step(0)
    assert e.value.code == -10
    x, y, z = ctx.download_positions()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(y)) and np.all(np.isfinite(z))
    # a non-finite upload is refused on the host
    bad = pos.copy()
    bad[3, 1] = np.nan
    with pytest.raises(S.SmdltError) as e:
        ctx.upload_positions(*xyz(bad))
    assert e.value.code == -1
    ctx.close()


def test_device_noise_moments_and_independence():
    # free draining with no forces and no flow: each step moves a bead by
    # sqrt(2 dt) z with z the device noise (Philox4x32-10 + Box-Muller)
    L, nb, nc, dt = 400.0, 100, 100, 1e-2
    ctx, p = ctx_for(0, L, nb, nc, prec=S.PREC_FP64, edot=0.0, dt=dt, kappas=0.0, epsilon=0.0,
                     noise_mode=S.NOISE_FREE_DRAINING)
    g = np.linspace(-150, 150, 22)
    pos = np.stack(np.meshgrid(g, g, g[:21], indexing="ij"), -1).reshape(-1, 3)[:nb * nc]  # 14 apart: no LJ
    ctx.upload_positions(*xyz(pos))
    zs = []
    for t in range(3):
        a = np.stack(ctx.download_positions(), 1)
        ctx.step(t)
        b = np.stack(ctx.download_positions(), 1)
        zs.append(((b - a) / np.sqrt(2 * dt)).reshape(-1))
    z = np.concatenate(zs)
    n = len(z)
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs((z ** 3).mean()) < 5 * np.sqrt(15 / n)
    assert abs((z ** 4).mean() - 3) < 5 * np.sqrt(96 / n)
    # components, neighbouring beads and consecutive steps are uncorrelated
    Z = zs[0].reshape(-1, 3)
    for c1, c2 in ((0, 1), (1, 2), (0, 2)):
        assert abs(np.corrcoef(Z[:, c1], Z[:, c2])[0, 1]) < 5 / np.sqrt(len(Z))
    assert abs(np.corrcoef(Z[:-1, 0], Z[1:, 0])[0, 1]) < 5 / np.sqrt(len(Z))
    assert abs(np.corrcoef(zs[0], zs[1])[0, 1]) < 5 / np.sqrt(len(zs[0]))
    # the histogram follows the normal density (chi^2 over 20 bins)
    from math import erf
    edges = np.linspace(-3, 3, 21)
    cdf = np.array([0.5 * (1 + erf(e / np.sqrt(2))) for e in edges])
    expct = n * np.diff(cdf)
    obs = np.histogram(z, edges)[0]
    chi2 = ((obs - expct) ** 2 / expct).sum()
    assert chi2 < 60  # 19 degrees of freedom; p ~ 3e-6
    ctx.close()


@pytest.mark.parametrize("graph", [False, True])
def test_run_after_strain_jump_rebuilds_lists(graph):
    """Steps that resume at a much later strain (a jump of the lattice) rebuild
    the neighbour lists: they equal the same steps of a fresh context holding
    the same positions, bit for bit (FP64)."""
    L, nb, nc = 30.0, 50, 4
    ctx, p = ctx_for(1, L, nb, nc, prec=S.PREC_FP64, edot=0.3, dt=5e-4)
    ctx.upload_positions(*xyz(random_walk(nb, nc, L, 10)))
    ctx.run(0, 3, use_graph=graph)
    mid = ctx.download_positions()
    ctx.run(4000, 3, use_graph=graph)
    fresh, _ = ctx_for(1, L, nb, nc, prec=S.PREC_FP64, edot=0.3, dt=5e-4)
    fresh.upload_positions(*mid)
    fresh.run(4000, 3, use_graph=graph)
    a, b = np.stack(ctx.download_positions(), 1), np.stack(fresh.download_positions(), 1)
    assert np.array_equal(a, b)
    ctx.close()
    fresh.close()


@pytest.mark.parametrize("mode", [S.NOISE_FREE_DRAINING, S.NOISE_TEA])
def test_run_in_other_noise_modes_equals_single_steps(mode):
    """smdlt_run in the free-draining and TEA modes steps eagerly; with device noise seeded by
    the step index it equals the same steps taken one by one."""
    L, nb, nc = 30.0, 50, 4
    out = []
    for use_run in (False, True):
        ctx, p = ctx_for(1, L, nb, nc, prec=S.PREC_FP64, edot=0.3, dt=5e-4, noise_mode=mode)
        if mode == S.NOISE_TEA:
            ctx.set_tea(0.4, np.full(3 * nb * nc, 0.8))
        rx, ry, rz = xyz(random_walk(nb, nc, L, 10))
        ctx.upload_positions(rx, ry, rz)
        if use_run:
            assert ctx.run(0, 5, use_graph=True) == 0
        else:
            for t in range(5):
                assert ctx.step(t) == 0
        out.append(np.stack(ctx.download_positions(), 1))
        ctx.close()
    assert np.array_equal(out[0], out[1])  # cells are sorted, so list order and sums are fixed


def test_tea_step_without_coefficients_is_a_state_error():
    ctx, p = ctx_for(1, 30.0, 20, 2, prec=S.PREC_FP64, noise_mode=S.NOISE_TEA)
    rx, ry, rz = xyz(random_walk(20, 2, 30.0, 3))
    ctx.upload_positions(rx, ry, rz)
    with pytest.raises(S.SmdltError):
        ctx.step(0)


def test_batched_replicas_equal_single_contexts():
    specs = [(1, 26.0, 30, 2), (2, 30.0, 50, 3), (1, 22.0, 17, 5)]
    params = []
    poss = []
    for k, (flow, L, nb, nc) in enumerate(specs):
        params.append(S.default_params(nb=nb, nc=nc, L=L, flow_type=flow, edot=0.4, dt=1e-3,
                                       precision=S.PREC_FP64, rc_real=6.0, grid=(32, 64, 32), spline_order=6))
        poss.append(random_walk(nb, nc, L, 20 + k))
    big = S.Context(params)
    allp = np.concatenate(poss)
    big.upload_positions(*xyz(allp))
    big.remap(120)
    v = np.random.default_rng(0).normal(size=3 * len(allp))
    ub = big.mobility_apply(v)
    off = 0
    for k, prm in enumerate(params):
        one = S.Context(prm)
        one.upload_positions(*xyz(poss[k]))
        one.remap(120)
        n = prm.nb * prm.nc
        u1 = one.mobility_apply(v[3 * off:3 * (off + n)])
        assert rel(ub[3 * off:3 * (off + n)], u1) < 1e-12
        off += n
    # the per-replica (plain) path on the same context gives the same numbers
    big.select_replicas(1, 2)
    u_plain = big.mobility_apply(v)
    n0 = params[0].nb * params[0].nc
    n1 = params[1].nb * params[1].nc
    assert rel(u_plain[3 * n0:3 * (n0 + n1)], ub[3 * n0:3 * (n0 + n1)]) < 1e-12


def test_large_grid_replicas_equal_single_contexts():
    # four mixed-precision replicas on 192^3 grids (340 MB of grids in one launch
    # sequence, so the grid passes run in two chunks of at most 256 MB: the first
    # three replicas, then the fourth) give exactly the products of one context
    # per replica
    specs = [(1, 40.0, 50, 4), (2, 44.0, 40, 5), (1, 36.0, 60, 3), (0, 40.0, 30, 6)]
    params, poss = [], []
    for k, (flow, L, nb, nc) in enumerate(specs):
        params.append(S.default_params(nb=nb, nc=nc, L=L, flow_type=flow, edot=0.4, dt=1e-3,
                                       precision=S.PREC_MIXED, rc_real=8.0, grid=(192, 192, 192), spline_order=6))
        poss.append(random_walk(nb, nc, L, 40 + k))
    big = S.Context(params)
    allp = np.concatenate(poss)
    big.upload_positions(*xyz(allp))
    big.remap(60)
    v = np.random.default_rng(1).normal(size=3 * len(allp))
    big.set_profiling(True)
    ub = big.mobility_apply(v)
    st = big.stage_times()
    big.set_profiling(False)
    # each grid pass is timed once per product, over both chunks
    assert all(st[k][1] == 1 for k in ("k_fft_x(fwd)", "k_fft_y(fwd)", "k_fft_z_green", "k_fft_y(inv)",
                                        "k_fft_x(inv)", "k_mob_final"))
    off = 0
    for k, prm in enumerate(params):
        one = S.Context(prm)
        one.upload_positions(*xyz(poss[k]))
        one.remap(60)
        n = prm.nb * prm.nc
        u1 = one.mobility_apply(v[3 * off:3 * (off + n)])
        assert rel(ub[3 * off:3 * (off + n)], u1) < 1e-12
        off += n
        one.close()


# ----------------------------------------------------------------- MC
@pytest.mark.parametrize("flow,nq,b1,b2", [(1, 24, 1.0, 1.5), (2, 40, 0.5, 0.2), (1, 60, 3.0, 3.0)])
def test_mc_reference_sequence_bit_exact(flow, nq, b1, b2):
    L, nb, nc = 22.0, 30, 4
    ctx, p = ctx_for(flow, L, nb, nc, edot=0.3, lambda_d=5.0, lambda_b=3.59, barrier=b1, barrier2=b2)
    lat = O.lattice(flow, L, 0.3, 1e-3, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 12))
    ctx.upload_positions(rx, ry, rz)
    ctx.remap(40)
    rx, ry, rz = oracle_remap(lat, 40, rx, ry, rz)
    track = np.random.default_rng(3).choice(nb * nc, nq, replace=False).astype(np.int32)
    ctx.upload_charges([nq], track)
    mc = O.ChargeHop(nb, nc, track)
    rng = O.Ran1(830062920)
    for sweep in range(4):
        u = rng.uniforms(3 * nq)
        cons, log = ctx.mc_sweep(S.MC_REFERENCE, u[None, :])
        st, cur, olog, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, b1, b2, u)
        tr, old = ctx.download_charges()
        assert cons[0] == cur
        assert np.array_equal(tr, mc.track)
        # per-trial records: from, to (skips: -1), kind, accepted, draws
        assert np.array_equal(log[:, 1], olog[:, 1])
        assert np.array_equal(log[:, 4], olog[:, 4]) and np.array_equal(log[:, 5], olog[:, 5])
        keep = olog[:, 3] >= 0
        assert np.array_equal(log[keep, 2], olog[keep, 2]) and np.array_equal(log[keep, 3], olog[keep, 3])


@pytest.mark.parametrize("case", range(8))
def test_mc_hand_worked_branches_match_oracle(case):
    """The chargeHop branches pinned by hand in the oracle tests (close-bead left
    and right hops, pops with barrier2, blocked pops and neighbours that leave
    the flag set, chain ends, rejections) give the same record on the GPU."""
    from test_oracle import HOP_CASES, hop_geometry
    track, u, b1, b2, expect = HOP_CASES[case]
    lat, rx, ry, rz = hop_geometry()
    ctx, p = ctx_for(0, 40.0, 6, 1, edot=0.0, lambda_d=5.0, lambda_b=3.59, barrier=b1, barrier2=b2)
    ctx.upload_positions(rx, ry, rz)
    ctx.remap(0)
    ctx.upload_charges([len(track)], np.array(track, np.int32))
    uu = np.array(u, np.float32)
    cons, log = ctx.mc_sweep(S.MC_REFERENCE, uu[None, :])
    mc = O.ChargeHop(6, 1, track)
    st, cur, olog, e0 = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, b1, b2, uu)
    assert cons[0] == cur
    assert np.array_equal(log[:, 1], olog[:, 1]) and np.array_equal(log[:, 4:], olog[:, 4:])
    keep = olog[:, 3] >= 0
    assert np.array_equal(log[keep, 2:4], olog[keep, 2:4]) and np.array_equal(log[:, 3] >= 0, keep)
    tr, old = ctx.download_charges()
    assert np.array_equal(tr, mc.track)


def test_mc_energy_matches_oracle():
    L, nb, nc = 25.0, 40, 3
    ctx, p = ctx_for(2, L, nb, nc, lambda_d=3.0)
    lat = O.lattice(2, L, 0.3, 1e-3, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 13))
    ctx.upload_positions(rx, ry, rz)
    ctx.remap(9)
    rx, ry, rz = oracle_remap(lat, 9, rx, ry, rz)
    track = np.arange(0, 120, 3, dtype=np.int32)
    sign = np.where(np.arange(len(track)) % 3 == 0, -1, 1).astype(np.int32)
    ctx.upload_charges([len(track)], track, sign)
    q = np.zeros(120, np.int32)
    q[track] = sign
    e = ctx.mc_energy()[0]
    assert e == pytest.approx(O.coulomb_energy(lat, rx, ry, rz, q, 3.59, 3.0), rel=1e-12)


def _raw_sweep(ctx, mode, u, ntot):
    """smdlt_mc_sweep called directly, so that the consumed counts and records
    of a sweep that ends with an error code can be read."""
    import ctypes as C
    u = np.ascontiguousarray(u, np.float32)
    cons = np.zeros(ctx.nrep, np.int32)
    log = np.zeros(6 * ntot, np.int32)
    code = S.lib().smdlt_mc_sweep(ctx.h, mode, u.ctypes.data_as(C.POINTER(C.c_float)), u.shape[-1],
                                  cons.ctypes.data_as(C.POINTER(C.c_int)), log.ctypes.data_as(C.POINTER(C.c_int)))
    return code, cons, log.reshape(-1, 6)


@pytest.mark.parametrize("flow,b1,b2", [(1, 0.4, 0.9), (2, 2.0, 0.0)])
def test_mc_reference_many_rounds_bit_exact(flow, b1, b2):
    # 200 charges: the sweep's blocks of 64 and the folds between them
    L, nb, nc, nq = 30.0, 50, 8, 200
    ctx, p = ctx_for(flow, L, nb, nc, edot=0.3, lambda_d=5.0, lambda_b=3.59, barrier=b1, barrier2=b2)
    lat = O.lattice(flow, L, 0.3, 1e-3, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 31))
    ctx.upload_positions(rx, ry, rz)
    ctx.remap(7)
    rx, ry, rz = oracle_remap(lat, 7, rx, ry, rz)
    track = np.random.default_rng(5).choice(nb * nc, nq, replace=False).astype(np.int32)
    ctx.upload_charges([nq], track)
    mc = O.ChargeHop(nb, nc, track)
    rng = O.Ran1(4242)
    acc = 0
    for sweep in range(3):
        u = rng.uniforms(2 * nq + 8)
        cons, log = ctx.mc_sweep(S.MC_REFERENCE, u[None, :])
        st, cur, olog, _ = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, b1, b2, u)
        assert st == 0 and cons[0] == cur
        assert np.array_equal(ctx.download_charges()[0], mc.track)
        assert np.array_equal(log[:, 1], olog[:, 1]) and np.array_equal(log[:, 4:], olog[:, 4:])
        keep = olog[:, 3] >= 0
        assert np.array_equal(log[keep, 2:4], olog[keep, 2:4])
        acc += int(olog[:, 4].sum())
    assert acc > 50
    ctx.close()


def test_mc_reference_stream_exhaustion_matches_oracle():
    # the uniforms run out in the third block: the charges before it are decided
    # as the oracle decides them, the count consumed is the whole stream and the
    # sweep reports SMDLT_E_MC_STREAM
    L, nb, nc, nq = 30.0, 50, 8, 200
    ctx, p = ctx_for(1, L, nb, nc, edot=0.3, lambda_d=5.0, lambda_b=3.59, barrier=0.6, barrier2=0.9)
    lat = O.lattice(1, L, 0.3, 1e-3, 1)
    rx, ry, rz = xyz(random_walk(nb, nc, L, 32))
    ctx.upload_positions(rx, ry, rz)
    ctx.remap(3)
    rx, ry, rz = oracle_remap(lat, 3, rx, ry, rz)
    track = np.random.default_rng(6).choice(nb * nc, nq, replace=False).astype(np.int32)
    ctx.upload_charges([nq], track)
    mc = O.ChargeHop(nb, nc, track)
    u = O.Ran1(99).uniforms(190)
    code, cons, log = _raw_sweep(ctx, S.MC_REFERENCE, u[None, :], nq)
    st, cur, olog, _ = mc.sweep(lat, rx, ry, rz, 3.59, 5.0, 0.6, 0.9, u)
    assert code == -9 and st == -2 and cons[0] == cur == 190
    assert np.array_equal(ctx.download_charges()[0], mc.track)
    done = int(np.nonzero(olog[:, 5])[0].max()) + 1  # charges the oracle recorded
    assert done > 128
    assert np.array_equal(log[:done, 1], olog[:done, 1]) and np.array_equal(log[:done, 4:], olog[:done, 4:])
    ctx.close()
