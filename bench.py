#!/usr/bin/env python3
"""Throughput of the Brownian-dynamics timestep on one MI355X, in bead-steps/s.

Workload (the metric's configuration): semidilute polyelectrolyte under shear,
40 chains x 200 beads (8000 beads), box L = 80a, Wi = 50 (edot = 50/175, the
notebooks' Wi = 175 edot), dt = 5e-4, 50 % of the beads charged +-1 with zero
net charge, Debye length 5a, Bjerrum length 1a, Hookean springs of the reference
defaults.  One step = remap + neighbour check + forces + device noise +
M.F + Lanczos M^(1/2) z (tol 1e-6, <= 30 iterations) + position update,
mixed precision (FP32 pair/grid/FFT sums, FP64 Lanczos), replayed as a hipGraph.

Also reported: the same workload on the plain per-replica FP64 path (no graph,
host-checked Lanczos exit), the M.F error of the mixed path against a tight
FP64 evaluation, per-kernel device times (HIP events on the context stream),
the roofline of the dominant kernel and the CPU restatement of the reference
algorithm (dense Ewald M every 100 steps + TEA matvec) timed on the host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "smdlt-ca-chrg-flow_amd"))
import smdlt_abi as S  # noqa: E402

HBM_PEAK_GBS = 8000.0


def chains(nb, nc, L, seed, bond=2.05, min_sep=2.0, tries=1000):
    """Seeded self-avoiding random walks like initChains (S/main26.c:420-470): bond
    2.05a, a new bead is redrawn while it lies within min_sep of any earlier bead
    (minimum image in the cubic box); positions wrapped into the box."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((nb * nc, 3))
    k = 0
    for c in range(nc):
        for j in range(nb):
            for _ in range(tries):
                if j == 0:
                    p = rng.uniform(-L / 2, L / 2, 3)
                else:
                    th = rng.uniform(0, 2 * np.pi)
                    ph = np.arccos(rng.uniform(-1, 1))
                    p = pos[k - 1] + bond * np.array([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)])
                p = p - L * np.round(p / L)
                if k == 0:
                    break
                d = pos[:k] - p
                d -= L * np.round(d / L)
                if np.min(np.einsum("ij,ij->i", d, d)) >= min_sep * min_sep:
                    break
            pos[k] = p
            k += 1
    return pos


def c2_params(prec):
    return S.default_params(nb=200, nc=40, L=80.0, flow_type=S.FLOW_SHEAR, edot=50.0 / 175.0, dt=5e-4,
                            spring_fene=0, kappas=200.0, qmax=2.0, epsilon=0.1, lambda_b=1.0, lambda_d=5.0,
                            precision=prec, lanczos_tol=1e-6, lanczos_maxit=30, seed=0x5EED0000)


def c2_charges(n, seed=7):
    rng = np.random.default_rng(seed)
    track = rng.choice(n, n // 2, replace=False).astype(np.int32)
    sign = np.where(np.arange(len(track)) < len(track) // 2, 1, -1).astype(np.int32)
    return track, sign


# Algorithmic HBM bytes per launch of each mobility kernel (one M.v):
# bead arrays are read/written once, grids once per pass (FP32 mixed path).
def kernel_bytes(n, G, nbr, order=8, esz=4):
    g = 3 * G * esz
    BW = 3 * order * esz
    return {
        "k_bead_prep": n * (24 + 24 + 12 + BW) ,
        "k_spread_gather": n * (12 + 24 + BW) + g,
        "k_fft_x(fwd)": 2 * g, "k_fft_y(fwd)": 2 * g, "k_fft_z_green": 2 * g,
        "k_fft_y(inv)": 2 * g, "k_fft_x(inv)": 2 * g,
        "k_mob_final": n * (24 + 24 + 24) + 4 * nbr + g,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "replicas shard one per GPU by design but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    pos = chains(200, 40, 80.0, 1)
    n = len(pos)
    track, sign = c2_charges(n)
    ctx = S.Context(c2_params(S.PREC_MIXED))
    cfg = ctx.mobility_config()
    ctx.upload_positions(pos[:, 0], pos[:, 1], pos[:, 2])
    ctx.upload_charges([len(track)], track, sign)
    print(f"[bench] mixed context ready: {cfg}", file=sys.stderr, flush=True)
    ctx.run(0, args.warmup, use_graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run(args.warmup, args.steps, use_graph=True)
    ctx.sync()
    el = time.perf_counter() - t0
    value = n * args.steps / el
    iters = int(ctx.last_iters()[0])
    print(f"[bench] timed {args.steps} steps in {el:.3f}s", file=sys.stderr, flush=True)

    # per-kernel device times on an eager (profiled) pass
    ctx.set_profiling(True)
    nprof = 3
    ctx.run(args.warmup + args.steps, nprof, use_graph=False)
    st = ctx.stage_times()
    ctx.set_profiling(False)
    G = cfg["grid"][0] * cfg["grid"][1] * cfg["grid"][2]
    nbr = n * 120  # mobility-list entries (estimated from the density within rc+skin)
    kb = kernel_bytes(n, G, nbr, cfg["order"])
    kern = {k: v for k, v in st.items() if k.startswith("k_") and v[1] > 0}
    dom = max(kern, key=lambda k: kern[k][0])
    avg_ms = kern[dom][0] / kern[dom][1]
    achieved = kb[dom] / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "c2_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get(dom)

    # M.F accuracy of the mixed path against a tight FP64 evaluation
    rng = np.random.default_rng(3)
    v = rng.normal(size=3 * n)
    p64 = c2_params(S.PREC_FP64)
    p64.mobility_tol = 1e-9
    ctx_t = S.Context(p64)
    ctx_t.upload_positions(*ctx.download_positions())
    ctx.update_lattice(args.warmup + args.steps + nprof)
    ctx_t.update_lattice(args.warmup + args.steps + nprof)
    u_mix = ctx.mobility_apply(v)
    u_ref = ctx_t.mobility_apply(v)
    mf_err = float(np.linalg.norm(u_mix - u_ref) / np.linalg.norm(u_ref))
    ctx_t.close()

    plain = None
    if not args.no_plain:
        c64 = S.Context(c2_params(S.PREC_FP64))
        c64.upload_positions(pos[:, 0], pos[:, 1], pos[:, 2])
        c64.upload_charges([len(track)], track, sign)
        c64.step(0)
        c64.sync()
        t1 = time.perf_counter()
        for k in range(args.plain_steps):
            c64.step(1 + k)
        c64.sync()
        plain = n * args.plain_steps / (time.perf_counter() - t1)
        c64.close()
        print(f"[bench] plain fp64 path {plain:.4g} bead-steps/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the checker's CPU restatement, timed as the baseline
        p = dict(nb=200, nc=40, flow=1, L=80.0, edot=50.0 / 175.0, dt=5e-4, kappas=200.0, qmax=2.0, epsilon=0.1,
                 lambda_b=1.0, lambda_d=5.0, barrier=3.0, barrier2=3.0, spring_fene=0, nq=0, mstep=0, srpy=100,
                 nt_avg=8, beta=0.5, seed=830062920)
        nsteps = 3
        elc, stc, _, thi = O.ref_run(p, pos[:, 0], pos[:, 1], pos[:, 2], nsteps, 0)
        per_step = (elc - thi) / nsteps + thi / 100.0
        cpu = {"value": n / per_step, "unit": "bead-steps/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"8000 beads: one dense Ewald M build (calcHI restated, far-pair table lookups) "
                         f"+ {nsteps} TEA steps; build time amortised over srpy=100 steps "
                         f"(build {thi:.2f}s, step {(elc - thi) / nsteps:.3f}s)"}
        print(f"[bench] cpu baseline {cpu['value']:.4g} bead-steps/s", file=sys.stderr, flush=True)

    out = {
        "metric": "bead-steps/s at Lanczos tol 1e-6, batched/learned vs plain GPU; % HBM roofline",
        "value": value, "unit": "bead-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pair/grid/FFT + fp64 Lanczos", "data": "synthetic",
        "config": {"workload": "semidilute polyelectrolyte under shear: 40x200 beads, L=80, Wi=50, dt=5e-4, "
                               "50% charged, lambda_D=5, lambda_B=1",
                   "grid": list(cfg["grid"]), "spline_order": cfg["order"], "rc": cfg["rc"],
                   "ewald_alpha": round(cfg["alpha"], 5), "lanczos_iters": iters, "parallelism": "1 GPU"},
        "mf_rel_err_vs_fp64": mf_err,
        "plain_fp64_value": plain,
        "speedup_vs_plain": (value / plain) if plain else None,
        "kernel_ms_avg": {k: v[0] / v[1] for k, v in kern.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch": kb[dom]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
