"""A compiled C99 consumer of the C ABI: tests/c_consumer/glue_run.c is the
glue a maintainer would add to the reference (gpu_init, the step, chargeHop fed
by the reference's own ran1 stream) around the time loop of S/main26.c:65-159.

The CPU test compiles and links it against include/smdlt_hip.h and the built
library. The GPU test runs it and replays the run in the oracle from the
program's own states: the ran1 stream (initChains, initial charges, getNoise's
gasdev deviates, chargeHop's uniforms and the advance by what a sweep consumed),
every step against the dense-matrix update (remap, bonds, LJ, Coulomb, dense
Ewald M, Jacobi square root; tolerance 1e-5 relative on the displacement) and
every sweep's records and charge placement bit for bit."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from helpers import O, ROOT, S

SRC = os.path.join(ROOT, "tests", "c_consumer", "glue_run.c")
# the run parameters written into glue_run.c
NB, NC, L, EDOT, DT, NQ, MSTEP, TMAX, SEED = 20, 3, 24.0, 0.5, 1e-3, 12, 2, 8, 830062920


def build_consumer(out):
    if not os.path.exists(S.LIB_PATH):
        pytest.skip("libsmdlt_hip.so is not built")
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    libdir = os.path.dirname(S.LIB_PATH)
    cmd = [gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-O2", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", str(out), "-L", libdir, "-lsmdlt_hip", f"-Wl,-rpath,{libdir}", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_c_consumer_compiles_and_links(tmp_path):
    exe = build_consumer(tmp_path / "glue_run")
    assert os.path.getsize(exe) > 0


def parse_run(path):
    lines = open(path).read().splitlines()
    n = NB * NC
    k = 0
    assert lines[k] == "init"
    init = np.array([[float(v) for v in lines[k + 1 + i].split()] for i in range(n)])
    track0 = np.array([int(lines[k + 1 + n + q]) for q in range(NQ)], np.int32)
    k += 1 + n + NQ
    steps = []
    while k < len(lines):
        tag, t = lines[k].split()
        assert tag == "step"
        pos = np.array([[float(v) for v in lines[k + 1 + i].split()] for i in range(n)])
        k += 1 + n
        mc = None
        if k < len(lines) and lines[k].startswith("mc "):
            consumed = int(lines[k].split()[1])
            log = np.array([[int(v) for v in lines[k + 1 + q].split()] for q in range(NQ)], np.int32)
            track = np.array([int(lines[k + 1 + NQ + q]) for q in range(NQ)], np.int32)
            mc = (consumed, log, track)
            k += 1 + 2 * NQ
        steps.append((int(t), pos, mc))
    return init, track0, steps


def oracle_step(lat, t, pos, q, z):
    """One updateHI step (S/main26.c:745-857) in the shear frame: remap, Hookean
    bonds (kappas 200, qmax 2), LJ, Coulomb, dense Ewald M, Jacobi M^(1/2)."""
    O.lattice_update(lat, t)
    wx, wy, wz = O.apply_pbc(lat, *[np.ascontiguousarray(pos[:, a]) for a in range(3)])
    nl, lst = O.verlet(lat, wx, wy, wz, 6.0)
    st, fb, _ = O.bond_force(lat, NB, NC, wx, wy, wz, 0, 200.0, 2.0)
    assert st == 0
    fl, _ = O.lj_force(lat, wx, wy, wz, nl, lst, 25.0, 0.1, 0)
    fc = O.coulomb_force(lat, wx, wy, wz, q, 3.59, 5.0)
    F = (fb + fl + fc).reshape(-1)
    w = np.stack([wx, wy, wz], 1)
    M = O.ewald_dense(O.cell_of(lat), w, 0.5)
    ys, _, _, _ = O.sym_sqrt_apply(M, z)
    d = (DT * (M @ F) + np.sqrt(2 * DT) * ys).reshape(-1, 3)
    d[:, 0] += DT * EDOT * wy
    return w, d


@pytest.mark.gpu
def test_c_consumer_run_matches_oracle(tmp_path):
    exe = build_consumer(tmp_path / "glue_run")
    out = tmp_path / "run.txt"
    print("C consumer: running glue_run", flush=True)
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    init, track0, steps = parse_run(out)
    assert len(steps) == TMAX
    n = NB * NC
    rng = O.Ran1(SEED)
    rx, ry, rz = rng.init_chains(NB, NC, L)
    assert np.array_equal(init, np.stack([rx, ry, rz], 1))
    track = rng.init_charges(n, NQ)
    assert np.array_equal(track0, track)
    mc = O.ChargeHop(NB, NC, track)
    lat = O.lattice(1, L, EDOT, DT, 1)
    prev = init
    worst, sweeps, accepted = 0.0, 0, 0
    for t, pos, rec in steps:
        z = rng.gauss(3 * n)
        q = np.zeros(n, np.int32)
        q[mc.track] = 1
        w, d = oracle_step(lat, t, prev, q, z)
        e = np.linalg.norm((pos - w) - d) / np.linalg.norm(d)
        worst = max(worst, e)
        assert e < 1e-5, (t, e)
        if rec is not None:
            consumed, log, gtrack = rec
            u = rng.copy().uniforms(3 * NQ + 8)
            gx, gy, gz = [np.ascontiguousarray(pos[:, a]) for a in range(3)]
            st, cur, olog, _ = mc.sweep(lat, gx, gy, gz, 3.59, 5.0, 1.0, 1.5, u)
            assert st == 0 and consumed == cur, t
            assert np.array_equal(log[:, 1], olog[:, 1]) and np.array_equal(log[:, 4:], olog[:, 4:]), t
            keep = olog[:, 3] >= 0
            assert np.array_equal(log[keep, 2:4], olog[keep, 2:4]), t
            assert np.array_equal(gtrack, mc.track), t
            rng.uniforms(consumed)
            sweeps += 1
            accepted += int(olog[:, 4].sum())
        prev = pos
    print(f"C consumer: {len(steps)} steps, worst step err {worst:.2e}; {sweeps} sweeps, {accepted} hops accepted")
