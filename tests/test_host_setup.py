"""The reference's run set-up on the host: its ran1/gasdev stream, initChains'
chains and charges, and the restart reader (S/Initialization.c:601-634,
S/main26.c:251-500,515-540).  The library's host functions are checked against
the oracle's restatement, which is pinned by the geometry the construction
guarantees; both run on the CPU."""
import numpy as np
import pytest
from helpers import O, S


def test_library_stream_equals_oracle_stream():
    for seed in (830062920, 981272199, -17, 0):
        lib, orc = S.RefStream(seed), O.Ran1(seed)
        assert np.array_equal(lib.uniforms(5000), orc.uniforms(5000))
        assert np.array_equal(lib.gauss(3001), orc.gauss(3001))  # odd count: the cached value carries over
        assert np.array_equal(lib.uniforms(7), orc.uniforms(7))


def chain_geometry_ok(rx, ry, rz, nb, nc, L):
    p = np.stack([rx, ry, rz], 1)
    assert np.all(np.abs(p) <= L / 2 + 1e-12)
    for k in range(nc):
        c = p[k * nb:(k + 1) * nb]
        d = np.diff(c, axis=0)
        d -= L * np.round(d / L)
        assert np.allclose(np.linalg.norm(d, axis=1), 2.05, atol=1e-12)
    for i in range(1, len(p)):
        d = p[i] - p[:i]
        d -= L * np.round(d / L)
        r2 = (d * d).sum(1)
        assert r2.min() >= (10.0 if i % nb == 0 else 4.0) - 1e-12, i


@pytest.mark.parametrize("nb,nc,L", [(30, 6, 30.0), (50, 1, 40.0), (12, 40, 28.0)])
def test_init_chains_geometry_and_equal_to_oracle(nb, nc, L):
    lib, orc = S.RefStream(830062920), O.Ran1(830062920)
    a = lib.init_chains(nb, nc, L)
    b = orc.init_chains(nb, nc, L)
    # equal up to libm: the two builds may reach sin/cos through different glibc
    # entry points (sin and cos of one angle can become one sincos call), an ulp apart
    assert all(np.abs(x - y).max() <= 1e-12 for x, y in zip(a, b))
    chain_geometry_ok(*b, nb, nc, L)
    # the first chain's first bead is the first three draws (one chain: the origin)
    u = O.Ran1(830062920).uniforms(3).astype(np.float64)
    first = np.array([b[0][0], b[1][0], b[2][0]])
    assert np.array_equal(first, np.zeros(3) if nc == 1 else u * L - L / 2)
    # the streams stay in step afterwards
    assert np.array_equal(lib.uniforms(10), orc.uniforms(10))


def test_init_charges_distinct_and_equal_to_oracle():
    lib, orc = S.RefStream(981272199), O.Ran1(981272199)
    t1, t2 = lib.init_charges(400, 150), orc.init_charges(400, 150)
    assert np.array_equal(t1, t2)
    assert len(set(t1.tolist())) == 150 and t1.min() >= 0 and t1.max() < 400
    # each charge is (int)(u * N) in float of the next draw not yet taken
    u = O.Ran1(981272199).uniforms(2000)
    seen, k = set(), 0
    for v in u:
        c = int(np.float32(v) * np.float32(400))
        if c not in seen:
            seen.add(c)
            assert t1[k] == c
            k += 1
            if k == 150:
                break


def test_restart_reader_keeps_the_last_frame(tmp_path):
    p = tmp_path / "R.xyz"
    S.write_restart(p, 5, [1.0, 2.0], [3.0, 4.0], [5.0, 6.0])
    with open(p, "a") as f:  # a second frame appended, as successive printRestart calls leave
        f.write("2\n17\nA -1.5 0.25 9.0\nA 7.0 8.0 -9.5\n")
    rx, ry, rz, t = S.read_restart(p, 2)
    assert t == 17 and np.array_equal(rx, [-1.5, 7.0]) and np.array_equal(ry, [0.25, 8.0])
    assert np.array_equal(rz, [9.0, -9.5])
    with pytest.raises(RuntimeError):
        S.read_restart(p, 3)  # a frame of another size
    with pytest.raises(RuntimeError):
        S.read_restart(tmp_path / "missing.xyz", 2)
