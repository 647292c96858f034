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
    ta = S.TeaAverages(nb, nc, sbins=4)
    rng = np.random.default_rng(0)
    samples = [(0.05, 0.3), (0.12, 0.4), (0.13, 0.5), (9.0, 0.2)]
    Cs = [rng.uniform(0.5, 1.0, 3 * nb * nc) for _ in samples]
    bins = [ta.add(st, 0.1, b, C) for (st, b), C in zip(samples, Cs)]
    assert bins == [0, 1, 1, 4]
    beta, Cv = ta.coefficients(1)
    assert beta == pytest.approx(0.45)
    want = (Cs[1].reshape(nc, -1).sum(0) + Cs[2].reshape(nc, -1).sum(0)) / (nc * 2)
    assert np.allclose(Cv, np.tile(want, nc))
    path = tmp_path / "DC.txt"
    ta.write_decomp(path, 1234)
    lines = path.read_text().splitlines()
    assert lines[0] == "1234" and len(lines) == 1 + 5 * (1 + nb)
    head = lines[1 + (1 + nb)].split()  # bin 1
    assert head[0] == "1" and head[1] == "2" and float(head[2]) == pytest.approx(0.45)
    row = [float(x) for x in lines[2 + (1 + nb)].split()]
    assert row[0] == 0 and np.allclose(row[1:], want[:3])
    assert lines[1 + 2 * (1 + nb)].split()[1] == "0"  # an empty bin is written as zeros


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
