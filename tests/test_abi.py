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
