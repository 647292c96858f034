"""CPU-side checks of the C ABI library: it loads and exports every symbol
include/smdlt_hip.h declares (no compute calls without a GPU)."""
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
