"""CPU-side checks of the C ABI library: it loads and exports every symbol
include/smdlt_hip.h declares (no compute calls without a GPU)."""
import numpy as np
import pytest
from helpers import S


def test_library_exports_every_declared_symbol():
    lib = S.lib()
    names = S.declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_strings_and_defaults():
    lib = S.lib()
    assert lib.smdlt_strerror(-3).decode() == "FENE overstretch"
    assert lib.smdlt_version().decode().startswith("smdlt")
    p = S.default_params()
    # reference defaults (S/Initialization.c:30-48, S/Input.txt)
    assert (p.nb, p.nc, p.barrier, p.lambda_d, p.kappas, p.qmax) == (50, 20, 3.0, 5.0, 200.0, 2.0)


def test_bad_arguments_rejected_without_device():
    import ctypes as C
    h = C.c_void_p()
    assert S.lib().smdlt_create(None, 1, 0, C.byref(h)) == -1
    assert S.lib().smdlt_num_beads(None, 0) == -1


def test_tea_averages_bins_and_decomp_layout(tmp_path):
    # CATEA's bookkeeping (S/main26.c:1317-1381) and printTEA's text (S/main26.c:1807-1826):
    # samples land in floor(strain/interval) capped at sbins; C is averaged over chains and samples
    nb, nc = 3, 2
    ta = S.TeaAverages(nb, nc, sbins=4, strain_interval=0.1)
    rng = np.random.default_rng(0)
    samples = [(0.05, 0.3), (0.12, 0.4), (0.13, 0.5), (9.0, 0.2)]
    Cs = [rng.uniform(0.5, 1.0, 3 * nb * nc) for _ in samples]
    bpath = tmp_path / "B.txt"
    bins = [ta.add(st, b, C, t=100 * k, beta_path=str(bpath)) for k, ((st, b), C) in enumerate(zip(samples, Cs))]
    assert bins == [0, 1, 1, 4]
    # decomp/B: "t beta" then "i Cx Cy Cz" per bead, %lf
    blines = bpath.read_text().splitlines()
    assert len(blines) == 4 * (1 + nb * nc) and blines[0] == "0 0.300000"
    assert blines[1] == "0 " + " ".join(f"{x:.6f}" for x in Cs[0][:3])
    path = tmp_path / "DC.txt"
    ta.write_decomp(str(path), 1234)
    lines = path.read_text().splitlines()
    want = (Cs[1].reshape(nc, -1).sum(0) + Cs[2].reshape(nc, -1).sum(0)) / (nc * 2)
    assert lines[0] == "1234" and len(lines) == 1 + 5 * (1 + nb)
    head = lines[1 + (1 + nb)].split()  # bin 1
    assert head[0] == "1" and head[1] == "2" and float(head[2]) == pytest.approx(0.45)
    row = [float(x) for x in lines[2 + (1 + nb)].split()]
    assert row[0] == 0 and np.allclose(row[1:], want[:3])
    assert lines[1 + 2 * (1 + nb)].split()[1] == "0"  # an empty bin is written as zeros
    # the DC reader gives the written means back (to the 13 digits of %.12e)
    rb, rc, t = S.tea_read(str(path), nb, 4)
    assert t == 1234 and rb[1] == pytest.approx(0.45, rel=1e-12) and np.allclose(rc[1], want, rtol=1e-12)
    assert rb[2] == 0.0 and np.all(rc[2] == 0.0)
    # resetAverage: the same means, 0/0 for an empty bin, and the sums restart
    beta, Cb, cnt = ta.finish()
    assert list(cnt) == [1, 2, 0, 0, 1] and beta[1] == pytest.approx(0.45) and np.allclose(Cb[1], want)
    assert np.isnan(beta[2]) and np.all(np.isnan(Cb[2]))
    assert list(ta.finish()[2]) == [0] * 5
    # updateHI's coefficients at a strain: the bin's beta, C of bead i = entry i mod 3Nb
    bt, Ct, b = S.tea_coefficients(nb, nc, 4, 0.1, 0.129, beta, Cb)
    assert b == 1 and bt == beta[1] and np.array_equal(Ct, np.tile(Cb[1], nc))
    assert S.tea_coefficients(nb, nc, 4, 0.1, 77.0, beta, Cb)[2] == 4


def write_ch(path, disp):
    """calcCH lines for disp[step][charge][3] (the charge indices are not read back)."""
    with open(path, "w") as f:
        for t, row in enumerate(disp):
            for j, d in enumerate(row):
                f.write(f"{t} {j} {j} {j} {d[0]:.6f} {d[1]:.6f} {d[2]:.6f}\n")


def test_msd_hand_worked_layout(tmp_path):
    # two charges, three output steps after t = 0, MSDstart = 0, x displacements
    # a1..a3 = 1, 2, 4 and b1..b3 = 8, 16, 32.  The sort puts a at slots 1..3 and
    # b at 5..7 (slots 0 and 4 stay zero); the windows read series q from q*3:
    # q = 0 sees (0, a1, a2, a3), q = 1 sees (a3, 0, b1, b2).  Shift 1: windows
    # of one entry, j = 0..2: (0 + 1 + 4)/3 and (16 + 0 + 64)/3; shift 2: (1 + 9)/2
    # and (16 + 64)/2; shift 3: 9 and (4 + 8)^2; MSD = mean over the two series.
    disp = np.zeros((4, 2, 3))
    disp[1:, 0, 0] = [1, 2, 4]
    disp[1:, 1, 0] = [8, 16, 32]
    ch = tmp_path / "ch.txt"
    write_ch(ch, disp)
    out = tmp_path / "msd.txt"
    r = S.msd(str(ch), 2, 3, 0, 1e-3, 10, str(out))
    want = [(85 / 3) / 2, 45 / 2, 153 / 2]
    assert np.allclose(r[:, 0], [0.01, 0.02, 0.03]) and np.allclose(r[:, 1], want) and np.allclose(r[:, 2], want)
    assert np.all(r[:, 3:] == 0)
    assert out.read_text().splitlines()[0] == f"0.010000,{want[0]:.6f},{want[0]:.6f},0.000000,0.000000"


def test_msd_constant_velocities_closed_form(tmp_path):
    # every charge moves by its own constant vector per output step: with the
    # windows inside each series' block (MSDstart >= number of series), the MSD
    # at shift i is i^2 times the mean squared step
    rng = np.random.default_rng(3)
    nq, T = 30, 40
    v = np.round(rng.normal(size=(nq, 3)), 3)
    disp = np.zeros((T + 1, nq, 3))
    disp[1:] = v[None]
    ch = tmp_path / "ch.txt"
    write_ch(ch, disp)
    r = S.msd(str(ch), nq, T, 26, 5e-4, 100)
    i = np.arange(1, T - 26 + 1)
    m = (v[:25] ** 2).sum(1).mean()  # the first 25 charges are averaged
    assert np.allclose(r[:, 1], i ** 2 * m, rtol=1e-12) and np.allclose(r[:, 0], i * 5e-4 * 100)
    # calcMSDcom on one series: windows from the second entry, shifts 1..T-1
    com = tmp_path / "com.txt"
    with open(com, "w") as f:
        for t in range(T + 1):
            d = v[0] if t else np.zeros(3)
            f.write(f"{t} {d[0]:.6f} {d[1]:.6f} {d[2]:.6f}\n")
    rc = S.msd_com(str(com), T, 5e-4, 100)
    j = np.arange(1, T)
    assert len(rc) == T - 1 and np.allclose(rc[:, 1], j ** 2 * (v[0] ** 2).sum(), rtol=1e-12)


def test_output_writers_follow_the_reference_formats(tmp_path):
    # the fprintf layouts of printVisc (S/main26.c:1874, X/main14.c:1894), printExt
    # (S/main26.c:971), calcRg (S/main26.c:958), printTrajectory (S/main26.c:874-884)
    # and printRestart (S/main26.c:917-921); %lf prints six decimals
    v = tmp_path / "visc.txt"
    S.append_visc(v, 1000, [1.5, -2.0, 3.5], 1000, S.FLOW_SHEAR)
    S.append_visc(v, 2000, [1.5, -2.0, 3.5], 1000, S.FLOW_EXTENSION)
    assert v.read_text() == "1000 0.003500\n2000 0.001500 -0.002000\n"
    e = tmp_path / "ext.txt"
    S.append_ext(e, 5, 0.25, [18.0, 9.0], 10, 2.0)
    assert e.read_text() == "5 0.250000 0.900000 0.810000\n5 0.250000 0.450000 0.202500\n"
    g = tmp_path / "rg.txt"
    S.append_rg(g, 7, np.array([[1.0, 0.5, -0.25, 2.0, 0.0, 3.0]]))
    assert g.read_text() == "7 0 1.000000 0.500000 -0.250000 2.000000 0.000000 3.000000\n"
    x = tmp_path / "traj.xyz"
    S.append_xyz_frame(x, 12, [10.0, 0.0, 2.5, 10.0], 10.0, [1.0, 2.0, 3.0], [0.0, 0.5, 1.0], [-1.0, 0.0, 1.0],
                       [0, 1, -1])
    assert x.read_text().splitlines() == [
        "3",
        'Lattice="10.000000 0.000000 0.000000 2.500000 10.000000 0.000000 0.000000 0.000000 10.000000" '
        'Origin="-6.250000 -5.000000 -5.000000" Properties="" Time=12',
        "A 1.000000 0.000000 -1.000000",
        "B 2.000000 0.500000 0.000000",
    ]
    r = tmp_path / "R.xyz"
    S.write_restart(r, 99, [1.0, 2.0], [3.0, 4.0], [5.0, 6.0])
    assert r.read_text() == "2\n99\nA 1.000000 3.000000 5.000000\nA 2.000000 4.000000 6.000000\n"


def test_charge_log_writer_follows_calcch(tmp_path):
    # calcCH's "%lu %d %d %d %Lf %Lf %Lf" lines (S/main26.c:2463,2480)
    c = tmp_path / "ch.txt"
    S.append_ch(c, 2000, [4, 9], [5, 7], np.array([[0.5, -1.25, 0.0], [1e-7, 2.0, -3.5]]))
    assert c.read_text() == "2000 0 4 5 0.500000 -1.250000 0.000000\n2000 1 9 7 0.000000 2.000000 -3.500000\n"
