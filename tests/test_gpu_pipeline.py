"""The reference's two-iteration run on the device (S/main26.c:65-243): the
host setup (ran1 stream, initChains), iteration 0 with TEA sampling every
strain_sample steps (ewaldBin + CATEA, S/main26.c:128-137), the strain-bin
averages, decomp/B and decomp/DC files and resetAverage (S/main26.c:1317-1381,
1807-1839), the restart file (S/main26.c:902-924) and iteration 1 stepping
with the bin coefficients updateHI chooses at the strain (S/main26.c:757-764).
Each stage is checked against the oracle or against the files it wrote."""
import numpy as np
import pytest
from helpers import O, S

pytestmark = pytest.mark.gpu

NB, NC, L, EDOT, DT = 10, 3, 16.0, 0.5, 1e-3
SBINS, SINTERVAL, SAMPLE = 3, 0.002, 2


def xyz(p):
    return [np.ascontiguousarray(p[:, a]) for a in range(3)]


def params(mode):
    return S.default_params(nb=NB, nc=NC, L=L, flow_type=S.FLOW_SHEAR, edot=EDOT, dt=DT, precision=S.PREC_FP64,
                            kappas=200.0, pbc_compat=1, noise_mode=mode, mobility_tol=1e-7)


def test_tea_iterations_restart_and_host_setup(tmp_path):
    n = NB * NC
    seed = 830062920
    rs = S.RefStream(seed)
    rx, ry, rz = rs.init_chains(NB, NC, L)
    orng = O.Ran1(seed)
    ox, oy, oz = orng.init_chains(NB, NC, L)
    assert np.array_equal(rx, ox) and np.array_equal(ry, oy) and np.array_equal(rz, oz)

    # iteration 0: Lanczos steps on the reference's noise, TEA samples every SAMPLE steps
    ctx = S.Context(params(S.NOISE_LANCZOS))
    ctx.upload_positions(rx, ry, rz)
    ta = S.TeaAverages(NB, NC, sbins=SBINS, strain_interval=SINTERVAL)
    bpath = str(tmp_path / "B.txt")
    sums = {}
    nsteps = 18
    for t in range(nsteps):
        z = rs.gauss(3 * n)
        assert np.array_equal(z, orng.gauss(3 * n))
        ctx.step(t, z)
        if t % SAMPLE:
            continue
        beta, Cv, _ = ctx.tea_sample()
        gx, gy, gz = ctx.download_positions()
        lat = O.lattice(1, L, EDOT, DT, 1, t)
        e, rsum, _, _ = O.ewald_bin(lat, gx, gy, gz, NB, L ** 3)
        b, Co = O.catea(n, e, rsum)
        assert beta[0] == pytest.approx(b, rel=1e-10) and np.abs(Cv - Co).max() < 1e-10
        strain = t * DT * EDOT
        k = ta.add(strain, beta[0], Cv, t=t, beta_path=bpath)
        assert k == min(int(np.floor(strain / SINTERVAL)), SBINS)
        sb, sc, cnt = sums.get(k, (0.0, np.zeros(3 * NB), 0))
        sums[k] = (sb + beta[0], sc + Cv.reshape(NC, 3 * NB).sum(0), cnt + 1)
    ctx.tea_sample()  # the context stays usable after sampling
    gx, gy, gz = ctx.download_positions()
    ctx.close()

    # decomp/B: "t beta" then n lines "i Cx Cy Cz" per sample
    blines = open(bpath).read().splitlines()
    assert len(blines) == len(range(0, nsteps, SAMPLE)) * (1 + n)
    assert blines[0].split()[0] == "0" and blines[1 + n].split()[0] == str(SAMPLE)

    # printTEA's decomp/DC, read back; resetAverage's means
    dc = str(tmp_path / "DC.txt")
    ta.write_decomp(dc, nsteps - 1)
    rb, rc, rt = S.tea_read(dc, NB, SBINS)
    beta_b, C_b, counts = ta.finish()
    assert rt == nsteps - 1
    for k in range(SBINS + 1):
        if k in sums:
            sb, sc, cnt = sums[k]
            assert counts[k] == cnt
            assert beta_b[k] == pytest.approx(sb / cnt, rel=1e-13)
            assert np.allclose(C_b[k], sc / (NC * cnt), rtol=1e-13, atol=0)
            assert rb[k] == pytest.approx(beta_b[k], rel=1e-11)
            assert np.allclose(rc[k], C_b[k], rtol=1e-11, atol=1e-300)
        else:
            assert counts[k] == 0 and np.isnan(beta_b[k])
    assert set(sums) == {0, 1, 2, 3}

    # restart: printRestart's file read back (%lf keeps six decimals)
    rpath = tmp_path / "restart.txt"
    S.write_restart(rpath, nsteps - 1, gx, gy, gz)
    qx, qy, qz, qt = S.read_restart(rpath, n)
    assert qt == nsteps - 1 and np.abs(np.stack([qx - gx, qy - gy, qz - gz])).max() <= 5e-7

    # iteration 1: TEA steps with the coefficients of the bin at each step's strain
    ctx = S.Context(params(S.NOISE_TEA))
    ctx.upload_positions(qx, qy, qz)
    prev = np.stack([qx, qy, qz], 1)
    for t in (3, 4, 13):
        strain = t * DT * EDOT
        beta, Cv, k = S.tea_coefficients(NB, NC, SBINS, SINTERVAL, strain, rb, rc)
        assert k == min(int(np.floor(strain / SINTERVAL)), SBINS)
        assert beta == rb[k] and np.array_equal(Cv, np.tile(rc[k], NC))
        ctx.set_tea(beta, Cv)
        z = rs.gauss(3 * n)
        ctx.step(t, z)
        pos = np.stack(ctx.download_positions(), 1)
        # oracle: remap, Hookean bonds, LJ, dense M and the TEA product of updateHI
        lat = O.lattice(1, L, EDOT, DT, 1)
        O.lattice_update(lat, t)
        wx, wy, wz = O.apply_pbc(lat, *xyz(prev))
        nl, lst = O.verlet(lat, wx, wy, wz, 6.0)
        st, fb, _ = O.bond_force(lat, NB, NC, wx, wy, wz, 0, 200.0, 2.0)
        fl, _ = O.lj_force(lat, wx, wy, wz, nl, lst, 25.0, 0.1, 0)
        w = np.stack([wx, wy, wz], 1)
        M = O.ewald_dense(O.cell_of(lat), w, 0.5)
        d = (DT * (M @ (fb + fl).reshape(-1)) + np.sqrt(2 * DT) * O.tea_apply(M, beta, Cv, z)).reshape(-1, 3)
        d[:, 0] += DT * EDOT * wy
        err = np.linalg.norm((pos - w) - d) / np.linalg.norm(d)
        print(f"TEA iteration 1, step {t} (strain bin {k}): displacement err {err:.2e}")
        assert err < 1e-5
        prev = pos
    ctx.close()
