#!/usr/bin/env python3
"""Throughput of the Brownian-dynamics timestep on one MI355X, in bead-steps/s.

Headline workload (`value`): the uneven replica ensemble, 512 independent
replicas (chain length U[50, 1000], chain count U{1..20}, box at volume
fraction 0.02, shear or planar extension with Wi U[1, 100]; 2.88 M beads),
advanced together: one launch per stage over all replicas, each step replayed
as hipGraphs, mixed precision (FP32 pair/grid/FFT sums with compensated
per-bead sums, FP64 Lanczos at tol 1e-6).  One step = remap + neighbour check
+ forces + device noise + M.F + Lanczos M^(1/2) z + position update.  Beside
it: the plain per-replica FP64 path (the same kernels in FP64, one replica per
launch sequence, eager, host-checked Lanczos exit) over all 512 replicas, the
roofline of the dominant kernel, and the CPU restatement of the reference
algorithm timed on the host over a seeded sample of the replicas.

Also reported (`semidilute`): the 8000-bead semidilute polyelectrolyte under
shear (40 x 200 beads, L = 80, Wi = 50, 4000 charges +-1), with its plain
FP64 path and the M.F error of the mixed path against a tight FP64 evaluation.
M.F and M^(1/2) z errors against the CPU oracle are measured by the tests.

Also reported (`charge_hopping`): 20 chains x 300 beads with 2400 charges
(hop radius 3): the sweep time of the reference chargeHop, of the exactly
balanced hop MC with uniform and with learned proposals, Brownian dynamics
with a learned-proposal sweep every 10 steps, and tau_int per sweep of the
Coulomb energy, of the charge configuration and of the charge distribution over
the chains, uniform against learned (64 replicas; the proposal trained on 3e5
uniform-proposal moves); then the learned proposal over hops within 8a against
the uniform proposal at 3a and at 8a (`far_hop`).
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
from smdlt_workloads import (c2_charges, c2_params, c4_params, c4_specs, chains, grow_chains,  # noqa: E402
                             grow_chains_grouped, to_lab_frame)

HBM_PEAK_GBS = 8000.0
# VALU issue: 256 CUs x 4 SIMDs, each SIMD taking 4 cycles per wave64 FP32 instruction, at 2.4 GHz
VALU_SIMDS, VALU_CLOCK_HZ, VALU_CYCLES_PER_INST = 1024, 2.4e9, 4
USE_GRAPH = True  # --eager: steps run eagerly (for tracing tools) instead of replaying hipGraphs


def tuner_summary(cfgs):
    """The autotuner's per-replica mobility settings (chosen for M.F error <= 1e-5;
    the configurations' nominal grid and cutoff are not used)."""
    g = [int(np.prod(c["grid"])) for c in cfgs]
    rc = [c["rc"] for c in cfgs]
    return {"spline_order": sorted({c["order"] for c in cfgs}), "rc": [min(rc), max(rc)],
            "grid_points": [min(g), max(g)], "ewald_alpha": [round(min(c["alpha"] for c in cfgs), 4),
                                                             round(max(c["alpha"] for c in cfgs), 4)]}


def ensemble_leg(nrep, steps, warmup, plain_steps):
    """The uneven replica ensemble, batched (graph-replayed, one launch per stage)
    against the plain per-replica FP64 path over the same replicas."""
    specs = c4_specs(nrep)
    t0 = time.perf_counter()
    pos = grow_chains_grouped([(sp["nb"], sp["nc"], sp["L"]) for sp in specs], 2)
    ctx = S.Context(c4_params(specs, S.PREC_MIXED))
    pos = to_lab_frame(ctx, specs, pos)
    allp = np.concatenate(pos)
    ntot = len(allp)
    print(f"[bench] ensemble: {nrep} replicas, {ntot} beads, chains grown in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    ctx.upload_positions(allp[:, 0], allp[:, 1], allp[:, 2])
    ctx.run(0, warmup, use_graph=USE_GRAPH)
    ctx.sync()
    t1 = time.perf_counter()
    code = ctx.run(warmup, steps, use_graph=USE_GRAPH)
    ctx.sync()
    el = time.perf_counter() - t1
    iters, status, _ = ctx.lanczos_status()
    value = ntot * steps / el
    print(f"[bench] ensemble {value:.4g} bead-steps/s ({1e3 * el / steps:.1f} ms/step, Lanczos iters "
          f"{int(iters.min())}..{int(iters.max())})", file=sys.stderr, flush=True)
    # per-kernel device times of one eager profiled step; algorithmic bytes summed
    # over the replicas; PMC traffic of the same kernel beside it
    ctx.set_profiling(True)
    ctx.run(warmup + steps, 1, use_graph=False)
    st = ctx.stage_times()
    ctx.set_profiling(False)
    cfgs = [ctx.mobility_config(r) for r in range(nrep)]
    G = sum(int(np.prod(c["grid"])) for c in cfgs)
    ent = sum(c["cluster_pairs"] or c["list_entries"] for c in cfgs)
    kb = kernel_bytes(ntot, G, ent, cfgs[0]["order"])
    kern = {k: v for k, v in st.items() if k in kb and v[1] > 0}
    dom = max(kern, key=lambda k: kern[k][0])
    avg_ms = kern[dom][0] / kern[dom][1]
    ach = kb[dom] / (avg_ms * 1e-3) / 1e9
    sym, traffic = pmc_traffic("ensemble", stage_kernels(dom, cfgs[0]["order"], bool(cfgs[0]["cluster_pairs"])))
    roof = {"bound": "hbm", "kernel": sym or dom, "stage": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch": kb[dom], "avg_launch_ms": avg_ms,
            "grid_points": G, "list_entries": ent, "valu_issue_frac": pmc_valu_frac("ensemble", sym, avg_ms),
            "note": "dominant kernel of one eager step (HIP events on the context stream); achieved = algorithmic "
                    "bytes per launch / average launch time; traffic = PMC FETCH+WRITE bytes per dispatch; "
                    "valu_issue_frac = PMC SQ_INSTS_VALU per dispatch x 4 cycles / (launch time x 1024 SIMDs x 2.4 GHz)"}
    kernel_ms = {k: v[0] / v[1] for k, v in kern.items()}
    ctx.close()
    out = {"value": value, "ms_per_step": 1e3 * el / steps, "beads": ntot, "replicas": nrep,
           "lanczos_iters_last_step": [int(iters.min()), int(iters.max())],
           "lanczos_converged": bool(code == 0 and np.all(status == 1)), "roofline": roof,
           "kernel_ms_avg": kernel_ms, "tuner": tuner_summary(cfgs), "specs": specs, "pos": pos}
    if plain_steps <= 0:
        return out
    # plain path: every replica on its own launch sequence, FP64, eager
    c64 = S.Context(c4_params(specs, S.PREC_FP64))
    c64.upload_positions(allp[:, 0], allp[:, 1], allp[:, 2])
    for r in range(nrep):  # a first step per replica builds its lists
        c64.select_replicas(r, r + 1)
        c64.step(0)
    c64.sync()
    t2 = time.perf_counter()
    for r in range(nrep):
        c64.select_replicas(r, r + 1)
        for k in range(plain_steps):
            c64.step(1 + k)
    c64.sync()
    plain = ntot * plain_steps / (time.perf_counter() - t2)
    c64.close()
    print(f"[bench] plain per-replica fp64 {plain:.4g} bead-steps/s over {nrep} replicas x {plain_steps} steps",
          file=sys.stderr, flush=True)
    out.update({"plain_fp64_value": plain, "plain_sample": f"all {nrep} replicas ({ntot} beads), {plain_steps} "
                                                           f"timed steps each after one list-building step",
                "speedup_vs_plain": value / plain})
    return out


def semidilute_leg(steps, warmup, plain_steps):
    """40 x 200 beads under shear with 4000 charges: graph-replayed mixed path, its
    roofline, the plain FP64 path and the mixed M.F error against tight FP64."""
    pos = chains(200, 40, 80.0, 1)
    n = len(pos)
    track, sign = c2_charges(n)
    ctx = S.Context(c2_params(S.PREC_MIXED))
    cfg = ctx.mobility_config()
    ctx.upload_positions(pos[:, 0], pos[:, 1], pos[:, 2])
    ctx.upload_charges([len(track)], track, sign)
    ctx.run(0, warmup, use_graph=USE_GRAPH)
    ctx.sync()
    t0 = time.perf_counter()
    code = ctx.run(warmup, steps, use_graph=USE_GRAPH)
    ctx.sync()
    el = time.perf_counter() - t0
    value = n * steps / el
    iters = int(ctx.last_iters()[0])
    ctx.set_profiling(True)
    nprof = 3
    ctx.run(warmup + steps, nprof, use_graph=False)
    st = ctx.stage_times()
    ctx.set_profiling(False)
    G = cfg["grid"][0] * cfg["grid"][1] * cfg["grid"][2]
    mc = ctx.mobility_config()
    nbr = mc["cluster_pairs"] or mc["list_entries"]
    kb = kernel_bytes(n, G, nbr, cfg["order"])
    kern = {k: v for k, v in st.items() if k.startswith("k_") and v[1] > 0}
    dom = max(kern, key=lambda k: kern[k][0])
    avg_ms = kern[dom][0] / kern[dom][1]
    achieved = kb[dom] / (avg_ms * 1e-3) / 1e9
    sym, traffic = pmc_traffic("semidilute", stage_kernels(dom, cfg["order"], bool(mc["cluster_pairs"])))
    rng = np.random.default_rng(3)
    v = rng.normal(size=3 * n)
    p64 = c2_params(S.PREC_FP64)
    p64.mobility_tol = 1e-9
    ctx_t = S.Context(p64)
    ctx_t.upload_positions(*ctx.download_positions())
    ctx.remap(warmup + steps + nprof)  # M is taken in the lattice frame
    ctx_t.remap(warmup + steps + nprof)
    u_mix = ctx.mobility_apply(v)
    u_ref = ctx_t.mobility_apply(v)
    mf_err = float(np.linalg.norm(u_mix - u_ref) / np.linalg.norm(u_ref))
    ctx_t.close()
    ctx.close()
    out = {"workload": "semidilute polyelectrolyte under shear: 40x200 beads, L=80, Wi=50, dt=5e-4, 50% charged "
                       "(+-1, zero net), lambda_D=5, lambda_B=1",
           "value": value, "unit": "bead-steps/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
           "lanczos_iters_last_step": iters, "lanczos_converged": code == 0,
           "tuner": {"grid": list(cfg["grid"]), "spline_order": cfg["order"], "rc": cfg["rc"],
                     "ewald_alpha": round(cfg["alpha"], 5),
                     "note": "the configuration names a 64^3 grid and rc = 5; the autotuner's choice holds "
                             "M.F error <= 1e-5 instead"},
           "mf_rel_err_vs_fp64": mf_err,
           "kernel_ms_avg": {k: v[0] / v[1] for k, v in kern.items()},
           "roofline": {"bound": "hbm", "kernel": sym or dom, "stage": dom, "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                        "bytes_per_launch": kb[dom], "avg_launch_ms": avg_ms, "list_entries": nbr,
                        "valu_issue_frac": pmc_valu_frac("semidilute", sym, avg_ms)}}
    if plain_steps > 0:
        c64 = S.Context(c2_params(S.PREC_FP64))
        c64.upload_positions(pos[:, 0], pos[:, 1], pos[:, 2])
        c64.upload_charges([len(track)], track, sign)
        c64.step(0)
        c64.sync()
        t1 = time.perf_counter()
        for k in range(plain_steps):
            c64.step(1 + k)
        c64.sync()
        plain = n * plain_steps / (time.perf_counter() - t1)
        c64.close()
        out.update({"plain_fp64_value": plain, "speedup_vs_plain": value / plain})
    print(f"[bench] semidilute {value:.4g} bead-steps/s ({out['ms_per_step']:.2f} ms/step)", file=sys.stderr,
          flush=True)
    return out


def cpu_baseline(specs, pos, budget_s=20.0):
    """The CPU restatement of the reference algorithm (dense Ewald M every srpy =
    100 steps with far-pair table lookups, TEA matvec per step; no charges, as in
    the ensemble) timed on the host over a seeded sample of the ensemble's
    replicas, until about budget_s of CPU work is done."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker's CPU restatement, timed as the baseline
    order = np.random.default_rng(17).permutation(len(specs))
    beads, secs, wall, used = 0, 0.0, 0.0, []
    for r in order:
        sp = specs[int(r)]
        if sp["nb"] * sp["nc"] > 6000:
            continue
        p = dict(nb=sp["nb"], nc=sp["nc"], flow=sp["flow"], L=sp["L"], edot=sp["edot"], dt=5e-4, kappas=200.0,
                 qmax=2.0, epsilon=0.1, lambda_b=1.0, lambda_d=5.0, barrier=3.0, barrier2=3.0, spring_fene=0, nq=0,
                 mstep=0, srpy=100, nt_avg=8, beta=0.5, seed=830062920)
        x = pos[int(r)]
        nsteps = 3
        elc, stc, _, thi = O.ref_run(p, x[:, 0], x[:, 1], x[:, 2], nsteps, 0)
        beads += len(x)
        secs += (elc - thi) / nsteps + thi / 100.0  # one step with the M build amortised
        wall += elc
        used.append(int(r))
        if (len(used) >= 3 and wall > budget_s) or len(used) >= 16:
            break
    return {"value": beads / secs, "unit": "bead-steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{len(used)} replicas of the ensemble (seeded, <= 6000 beads each, {beads} beads): "
                      f"dense Ewald M build (calcHI restated, far-pair table lookups) amortised over srpy=100 "
                      f"steps + TEA matvec per step, no charges; 3 steps each"}


def tau_int_multi(E, c=5.0, per_column_mean=False, groups=1):
    """Integrated autocorrelation time from the autocorrelation of the columns of E
    (each column centred on its own mean), summed within `groups` consecutive
    column blocks (one block per independent replica) and averaged over blocks,
    with Sokal's automatic window."""
    x = E - E.mean(0)
    n = len(x)
    f = np.fft.rfft(x, 2 * n, axis=0)
    acf = np.fft.irfft(f * np.conj(f), axis=0)[:n]
    acf = acf.reshape(n, groups, -1).sum(2).mean(1) if groups > 1 else acf.reshape(n, -1).mean(1)
    acf /= acf[0]
    tau = 1.0
    for W in range(1, n):
        tau = 1.0 + 2.0 * acf[1:W + 1].sum()
        if W >= c * tau:
            break
    return tau


def mh_study(ctx, R, nq, n, track, sign, train_moves, sweeps, burn, iters, hops_learned=True, tag="", nb_chain=None):
    """Train the proposal on the uniform chains' moves of ctx (R replicas), then run
    uniform and learned chains; tau_int of the energy, of the charge configuration
    (per-bead occupancy) and of the charge distribution over the chains (net charge
    per chain of nb_chain beads)."""
    nb_chain = nb_chain or n
    t0 = time.perf_counter()
    ctx.upload_charges([nq] * R, np.tile(track, R), np.tile(sign, R))
    F, Y = [], []
    while hops_learned and sum(len(v) for v in Y) < train_moves:
        ctx.mc_sweep(S.MC_MH_UNIFORM)
        for r in range(R):
            f, fr, to, a = ctx.mc_features(r)
            k = to >= 0
            F.append(f[k])
            Y.append(a[k])
    w, l0, l1 = None, None, None
    if hops_learned:
        F = np.concatenate(F)
        Y = np.concatenate(Y).astype(np.float32)
        t1 = time.perf_counter()
        w, l0, l1 = ctx.mlp_train(F, Y, S.mlp_init(0), iters=iters)
        ctx.mlp_set_weights(w)
        t2 = time.perf_counter()
        print(f"[bench] {tag}: {len(Y)} training moves in {t1 - t0:.1f}s, loss {l0:.3f} -> {l1:.3f} in "
              f"{t2 - t1:.1f}s", file=sys.stderr, flush=True)

    def chain(mode):
        ctx.upload_charges([nq] * R, np.tile(track, R), np.tile(sign, R))
        ctx.mc_record_features(False)  # production sweeps: no training records
        E = np.zeros((sweeps, R))
        occ = np.zeros((sweeps, R, nq), np.int32)
        acc = 0
        el = 0.0
        for s in range(sweeps):
            ctx.sync()
            ts = time.perf_counter()
            _, log = ctx.mc_sweep(mode)
            el += time.perf_counter() - ts
            acc += int(log[:, 4].sum())
            E[s] = ctx.mc_energy()
            tr, _ = ctx.download_charges()
            occ[s] = tr.reshape(R, nq)
        ctx.mc_record_features(True)
        sub = occ[burn:]
        ns = len(sub)
        O3 = np.zeros((ns, R, n), np.float32)
        for r in range(R):
            O3[np.arange(ns)[:, None], r, sub[:, r, :]] = sign[None, :]
        # net charge of each chain: the charge distribution over the chains
        Qc = O3.reshape(ns, R, n // nb_chain, nb_chain).sum(3)
        out = {"acceptance": acc / (sweeps * nq * R), "tau_E": float(tau_int_multi(E[burn:])),
               "tau_config": float(tau_int_multi(O3.reshape(ns, -1), per_column_mean=True, groups=R)),
               "tau_chain_charge": float(tau_int_multi(Qc.reshape(ns, -1), per_column_mean=True, groups=R)),
               "mean_E": float(E[burn:].mean()), "mean_E_se": float(E[burn:].mean(0).std(ddof=1) / np.sqrt(R)),
               "sweep_ms": 1e3 * el / sweeps}
        print(f"[bench] {tag} {'learned' if mode == S.MC_MH_LEARNED else 'uniform'}: {out}", file=sys.stderr,
              flush=True)
        return out
    res = {"uniform": chain(S.MC_MH_UNIFORM)}
    if hops_learned:
        res["learned"] = chain(S.MC_MH_LEARNED)
        res.update({"train_moves": int(len(Y)), "train_accept": float(Y.mean()), "loss": [l0, l1]})
    return res, w


def learned_leg(replicas=16, train_moves=1_000_000, sweeps=150, burn=30, iters=1000, strong=True, far_hop=8.0):
    """Charge hopping with the learned proposal.
    (1) 20 chains x 300 beads (L = 60, 40 % charged, lambda_B = 3.59, lambda_D = 5,
    hop radius 3a): bead-steps/s of Brownian dynamics with one learned-proposal MC
    sweep every 10 steps, and tau_int per sweep of the Coulomb energy and of the
    charge configuration, uniform against learned proposal (trained on 1e6
    uniform-proposal moves).  The same at hop radius far_hop: the learned proposal
    (trained on uniform moves at that radius) against the uniform proposal at 3a
    (the configuration's) and at far_hop.
    (2) A strongly coupled small system (4 x 30 beads, L = 18, 40 charges,
    lambda_B = 4) where the uniform proposal mixes slowly: uniform at 3a and 8a
    against the learned proposal at 8a."""
    nb, nc, L = 300, 20, 60.0
    n = nb * nc
    nq = int(0.4 * n)
    pos = grow_chains([(nb, nc, L)], 5)[0]
    rng = np.random.default_rng(11)
    track = np.sort(rng.choice(n, nq, replace=False)).astype(np.int32)
    sign = np.where(np.arange(nq) % 2 == 0, 1, -1).astype(np.int32)

    def params(r, hop, nb=nb, nc=nc, L=L, lb=3.59):
        return S.default_params(nb=nb, nc=nc, L=L, flow_type=S.FLOW_SHEAR, edot=50.0 / 175.0, dt=5e-4,
                                spring_fene=0, kappas=200.0, qmax=2.0, epsilon=0.1, lambda_b=lb, lambda_d=5.0,
                                barrier=0.0, barrier2=0.0, precision=S.PREC_MIXED, lanczos_tol=1e-6,
                                lanczos_maxit=128, seed=0x5EED0000 + r, hop_radius=hop)
    out = {"workload": f"{nc}x{nb} beads, L={L}, {nq} charges (+-1), lambda_B=3.59, lambda_D=5, shear Wi=50, "
                       f"hop radius 3",
           "mlp": "16-64-64-1 ReLU, BF16 weights, FP32 accumulation; trained by Adam on the GPU gradient"}
    bd = S.Context(params(0, 3.0))
    bd.upload_positions(pos[:, 0], pos[:, 1], pos[:, 2])
    bd.upload_charges([nq], track, sign)
    bd.run(0, 2, use_graph=USE_GRAPH)
    x, y, z = bd.download_positions()
    # sweep times of the three modes on one replica (reference chargeHop on
    # device uniforms; MH without training records)
    bd.mc_record_features(False)
    sweep_ms = {}
    for mode, name in ((S.MC_REFERENCE, "reference_chargehop"), (S.MC_MH_UNIFORM, "mh_uniform")):
        bd.upload_charges([nq], track, sign)
        bd.mc_sweep(mode)
        bd.sync()
        t0 = time.perf_counter()
        for _ in range(5):
            bd.mc_sweep(mode)
        bd.sync()
        sweep_ms[name] = 1e3 * (time.perf_counter() - t0) / 5
    bd.upload_charges([nq], track, sign)
    bd.mc_record_features(True)
    R = replicas
    ctx = S.Context([params(r, 3.0) for r in range(R)])
    ctx.upload_positions(np.tile(x, R), np.tile(y, R), np.tile(z, R))
    ctx.remap(2)
    hop3, w = mh_study(ctx, R, nq, n, track, sign, train_moves, sweeps, burn, iters, tag="20x300 hop 3", nb_chain=nb)
    ctx.close()
    far = None
    if far_hop:
        ctx = S.Context([params(r, far_hop) for r in range(R)])
        ctx.upload_positions(np.tile(x, R), np.tile(y, R), np.tile(z, R))
        ctx.remap(2)
        far, _ = mh_study(ctx, R, nq, n, track, sign, train_moves, sweeps, burn, iters, tag=f"20x300 hop {far_hop:g}",
                          nb_chain=nb)
        ctx.close()
    bd.mlp_set_weights(w)
    bd.mc_record_features(False)
    bd.mc_sweep(S.MC_MH_LEARNED)
    bd.sync()
    t0 = time.perf_counter()
    for _ in range(5):
        bd.mc_sweep(S.MC_MH_LEARNED)
    bd.sync()
    sweep_ms["mh_learned"] = 1e3 * (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    nblk = 3
    for b in range(nblk):
        bd.run(2 + 10 * b, 10, use_graph=USE_GRAPH)
        bd.mc_sweep(S.MC_MH_LEARNED)
    bd.sync()
    el = time.perf_counter() - t0
    bd.close()
    out.update({"bead_steps_per_s_bd_plus_mc": n * 10 * nblk / el, "mc_every": 10,
                "sweep_ms_one_replica": sweep_ms, "chains": hop3,
                "tau_E_ratio_uniform_over_learned": hop3["uniform"]["tau_E"] / hop3["learned"]["tau_E"],
                "tau_config_ratio_uniform_over_learned": hop3["uniform"]["tau_config"] / hop3["learned"]["tau_config"],
                "tau_chain_charge_ratio_uniform_over_learned":
                    hop3["uniform"]["tau_chain_charge"] / hop3["learned"]["tau_chain_charge"]})
    if far is not None:
        ratios = {}
        for key in ("tau_E", "tau_config", "tau_chain_charge"):
            ratios[key + "_ratio_uniform3_over_learned_far"] = hop3["uniform"][key] / far["learned"][key]
            ratios[key + "_ratio_same_radius_far"] = far["uniform"][key] / far["learned"][key]
        out["far_hop"] = {"hop_radius": far_hop, "chains": far, **ratios,
                          "note": "learned proposal over the hops within the far radius (trained on uniform moves "
                                  "there) against the configuration's uniform 3a proposal, and against the uniform "
                                  "proposal at the same radius"}
    print(f"[bench] charge hopping: sweeps {sweep_ms} ms, tau_E ratio {out['tau_E_ratio_uniform_over_learned']:.2f}"
          + (f", learned at {far_hop:g}a over uniform at 3a: tau_E {out['far_hop']['tau_E_ratio_uniform3_over_learned_far']:.2f}, "
             f"tau_chain_charge {out['far_hop']['tau_chain_charge_ratio_uniform3_over_learned_far']:.2f}" if far is not None else ""),
          file=sys.stderr, flush=True)
    if not strong:
        return out
    # strongly coupled small system
    nb2, nc2, L2, nq2, R2 = 30, 4, 18.0, 40, 64
    pos2 = grow_chains([(nb2, nc2, L2)], 6)[0]
    tr2 = np.sort(np.random.default_rng(3).choice(nb2 * nc2, nq2, replace=False)).astype(np.int32)
    sg2 = np.where(np.arange(nq2) % 2 == 0, 1, -1).astype(np.int32)
    small = {}
    for hop, learn in ((3.0, False), (8.0, True)):
        ctx = S.Context([params(r, hop, nb2, nc2, L2, 4.0) for r in range(R2)])
        ctx.upload_positions(np.tile(pos2[:, 0], R2), np.tile(pos2[:, 1], R2), np.tile(pos2[:, 2], R2))
        ctx.remap(0)
        small[hop], _ = mh_study(ctx, R2, nq2, nb2 * nc2, tr2, sg2, 250_000, 600, 100, 1500, learn,
                                 tag=f"4x30 lambda_B=4 hop {hop:g}")
        ctx.close()
    out.update({"strong_coupling": {"workload": "4x30 beads, L=18, 40 charges, lambda_B=4, lambda_D=5, 64 replicas",
                                    "hop3": small[3.0], "hop8": small[8.0],
                                    "tau_E_ratio_uniform3_over_learned8":
                                        small[3.0]["uniform"]["tau_E"] / small[8.0]["learned"]["tau_E"],
                                    "tau_E_ratio_same_radius8":
                                        small[8.0]["uniform"]["tau_E"] / small[8.0]["learned"]["tau_E"],
                                    "tau_config_ratio_uniform3_over_learned8":
                                        small[3.0]["uniform"]["tau_config"] / small[8.0]["learned"]["tau_config"]}})
    return out


def single_chain_leg(steps=200, warmup=5):
    """The reference's own small case: one FENE chain of 100 beads (b_max = 10,
    H = 30), L = 60, shear Wi = 10 (edot = 10/175), dt = 1e-3, no charges, FP64,
    Lanczos tol 1e-6 (<= 30 iterations), mobility settings auto-tuned to 1e-5."""
    p = S.default_params(nb=100, nc=1, L=60.0, flow_type=S.FLOW_SHEAR, edot=10.0 / 175.0, dt=1e-3, spring_fene=1,
                         kappas=30.0, qmax=10.0, precision=S.PREC_FP64, lanczos_tol=1e-6, lanczos_maxit=30,
                         seed=0x5EED0001)
    pos = grow_chains([(100, 1, 60.0)], 8)[0]
    ctx = S.Context(p)
    ctx.upload_positions(pos[:, 0], pos[:, 1], pos[:, 2])
    ctx.run(0, warmup, use_graph=USE_GRAPH)
    ctx.sync()
    t0 = time.perf_counter()
    code = ctx.run(warmup, steps, use_graph=USE_GRAPH)
    ctx.sync()
    el = time.perf_counter() - t0
    it, stt, ch = ctx.lanczos_status()
    cfg = ctx.mobility_config()
    ctx.close()
    return {"workload": "1 FENE chain x 100 beads, L=60, shear Wi=10, dt=1e-3, FP64", "value": 100 * steps / el,
            "unit": "bead-steps/s", "ms_per_step": 1e3 * el / steps, "lanczos_iters_last_step": int(it[0]),
            "grid": list(cfg["grid"]), "rc": cfg["rc"], "ewald_alpha": round(cfg["alpha"], 5), "return_code": code}


def extension_leg(steps=5, warmup=2):
    """Planar extensional flow with Kraynik-Reinelt remaps: 200 chains x 250 beads
    (50000 beads), L = 160, Wi = 20 (edot = 20/175), 30 % charged (+-1, zero net),
    Debye length 3, dt = 2e-4, mixed precision, Lanczos tol 1e-6."""
    nb, nc, L = 250, 200, 160.0
    n = nb * nc
    p = S.default_params(nb=nb, nc=nc, L=L, flow_type=S.FLOW_EXTENSION, edot=20.0 / 175.0, dt=2e-4, spring_fene=0,
                         kappas=200.0, qmax=2.0, epsilon=0.1, lambda_b=1.0, lambda_d=3.0, lj_ext_form=1,
                         precision=S.PREC_MIXED, lanczos_tol=1e-6, lanczos_maxit=128, seed=0x5EED0003)
    t0 = time.perf_counter()
    pos = grow_chains([(nb, nc, L)], 4)
    ctx = S.Context(p)
    pos = to_lab_frame(ctx, [{"flow": S.FLOW_EXTENSION}], pos)[0]
    rng = np.random.default_rng(13)
    nq = int(0.3 * n) // 2 * 2
    track = rng.choice(n, nq, replace=False).astype(np.int32)
    sign = np.where(np.arange(nq) < nq // 2, 1, -1).astype(np.int32)
    ctx.upload_positions(pos[:, 0], pos[:, 1], pos[:, 2])
    ctx.upload_charges([nq], track, sign)
    cfg = ctx.mobility_config()
    print(f"[bench] planar extension: {n} beads, {nq} charges, setup {time.perf_counter() - t0:.1f}s, {cfg}", file=sys.stderr,
          flush=True)
    ctx.run(0, warmup, use_graph=USE_GRAPH)
    ctx.sync()
    t1 = time.perf_counter()
    code = ctx.run(warmup, steps, use_graph=USE_GRAPH)
    ctx.sync()
    el = time.perf_counter() - t1
    it, stt, ch = ctx.lanczos_status()
    ctx.set_profiling(True)
    ctx.run(warmup + steps, 1, use_graph=False)
    stages = {k: round(v[0], 4) for k, v in ctx.stage_times().items() if v[1]}
    ctx.set_profiling(False)
    x, y, z = ctx.download_positions()
    ctx.close()
    out = {"workload": "planar extension (KRBC), 200x250 beads, L=160, Wi=20, 15000 charges, lambda_D=3, dt=2e-4",
           "value": n * steps / el, "unit": "bead-steps/s", "ms_per_step": 1e3 * el / steps,
           "lanczos_iters_last_step": int(it[0]), "lanczos_converged": bool(stt[0] == 1),
           "grid": list(cfg["grid"]), "rc": cfg["rc"], "ewald_alpha": round(cfg["alpha"], 5),
           "finite": bool(np.all(np.isfinite(x)) and np.all(np.isfinite(y)) and np.all(np.isfinite(z))),
           "stage_ms_one_eager_step": stages}
    print(f"[bench] planar extension: {out['value']:.4g} bead-steps/s ({out['ms_per_step']:.1f} ms/step)", file=sys.stderr,
          flush=True)
    return out


# Algorithmic HBM bytes per launch of each mobility kernel (one M.v, mixed
# precision): bead arrays read or written once, grids once per pass, the pair
# list once.  The real-space pass ("k_mob_final" stage) reads the FP32 position
# and vector copies (16 + 16 B), the FP64 input (24 B) for the self term, the
# stencil data (12 + 12p B), writes u (24 B), reads the grid once and its list:
# 4 B per cluster pair in mixed precision (4 B per bead pair in FP64).
def kernel_bytes(n, G, list_entries, order=8, esz=4):
    g = 3 * G * esz
    BW = 3 * order * esz
    return {
        "k_bead_prep": n * (24 + 16 + 12 + BW),
        "k_spread": n * (12 + 16 + BW) + g,
        "k_fft_x(fwd)": 2 * g, "k_fft_y(fwd)": 2 * g, "k_fft_z_green": 2 * g,
        "k_fft_y(inv)": 2 * g, "k_fft_x(inv)": 2 * g,
        "k_mob_final": n * (16 + 16 + 24 + 24 + 12 + BW) + 4 * list_entries + g,
    }


def stage_kernels(stage, order, cluster):
    """Kernel symbols a per-stage timing slot can hold: the last slot of the
    mobility product runs the cluster pass (mixed precision) or the bead pass."""
    if stage == "k_mob_final" and cluster:
        return [f"k_mob_cl<{order}, "]
    if stage == "k_spread":
        return [f"k_spread_mfma<{order}>", "k_spread_gather<"]
    return [stage, stage.split("(")[0]]


def pmc_valu_frac(section, sym, avg_ms):
    """Fraction of the VALU issue peak the kernel used: its SQ_INSTS_VALU per
    dispatch from the committed counter summary, 4 cycles per instruction, over
    the 1024 SIMDs for the measured average launch time; None when absent."""
    tf = os.path.join(ROOT, "profiles", "hbm_traffic_pmc.json")
    if not sym or not os.path.exists(tf):
        return None
    with open(tf) as fh:
        d = json.load(fh).get(section + "_valu_insts", {})
    n = d.get(sym)
    if n is None or avg_ms <= 0:
        return None
    return n * VALU_CYCLES_PER_INST / (avg_ms * 1e-3 * VALU_SIMDS * VALU_CLOCK_HZ)


def pmc_traffic(section, names):
    """HBM bytes per dispatch of the first kernel of the committed FETCH_SIZE +
    WRITE_SIZE summary (scaled by the copy calibration) named by `names`: an exact
    name, or a template prefix ending in ", "; (None, None) when none is there."""
    tf = os.path.join(ROOT, "profiles", "hbm_traffic_pmc.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            d = json.load(fh).get(section, {})
        for n in names:
            for k, v in d.items():
                if k == n or (n.endswith((", ", "<")) and k.startswith(n)):
                    return k, v
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed steps of the replica ensemble")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--semidilute-steps", type=int, default=30)
    ap.add_argument("--hop-sweeps", type=int, default=100,
                    help="sweeps per chain of the charge-hopping workload (burn-in: a tenth more)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--workload", choices=["all", "ensemble", "semidilute", "hop", "learned", "extension", "single-chain"], default="all",
                    help="ensemble: only the replica ensemble; semidilute: only the 8000-bead shear line; hop: only the "
                         "charge-hopping line; learned: charge hopping with the strongly coupled study")
    ap.add_argument("--eager", action="store_true", help="run the timed steps eagerly instead of as hipGraphs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device-to-device three times (HBM counter calibration)")
    args = ap.parse_args()
    global USE_GRAPH
    USE_GRAPH = not args.eager
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "replicas shard one per GPU by design but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        a = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}))
        return
    if args.workload == "single-chain":
        print(json.dumps({"single_chain": single_chain_leg()}))
        return
    if args.workload == "extension":
        print(json.dumps({"planar_extension": extension_leg()}))
        return
    if args.workload == "learned":
        print(json.dumps({"learned": learned_leg()}, default=float))
        return
    if args.workload == "hop":
        burn = max(20, args.hop_sweeps // 10)
        print(json.dumps({"charge_hopping": learned_leg(replicas=64, train_moves=300_000, sweeps=args.hop_sweeps,
                                                        burn=burn, iters=600, strong=False)}, default=float))
        return
    if args.workload == "semidilute":
        print(json.dumps({"semidilute": semidilute_leg(args.semidilute_steps, 3, 0 if args.no_plain else 3)}))
        return

    ens = ensemble_leg(args.replicas, args.steps, args.warmup, 0 if args.no_plain else args.plain_steps)
    specs, pos = ens.pop("specs"), ens.pop("pos")
    semi = None
    if args.workload == "all":
        semi = semidilute_leg(args.semidilute_steps, 3, 0 if args.no_plain else 3)
    hop = None
    if args.workload == "all":
        hop = learned_leg(replicas=64, train_moves=300_000, sweeps=100, burn=20, iters=600, strong=False)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(specs, pos)
        print(f"[bench] cpu baseline {cpu['value']:.4g} bead-steps/s", file=sys.stderr, flush=True)
    out = {
        "metric": "bead-steps/s at Lanczos tol 1e-6, batched/learned vs plain GPU; % HBM roofline",
        "value": ens["value"], "unit": "bead-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ens["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pair/grid/FFT + fp64 Lanczos", "data": "synthetic",
        "config": {"workload": f"batched ensemble: {args.replicas} uneven replicas (chain length U[50,1000], chain "
                               f"count U{{1..20}}, volume fraction 0.02, shear or planar extension, Wi U[1,100], "
                               f"dt=5e-4), no charges, one launch per stage, steps replayed as hipGraphs",
                   "replicas": args.replicas, "beads": ens["beads"], "lanczos_tol": 1e-6,
                   "lanczos_iters_last_step": ens["lanczos_iters_last_step"],
                   "lanczos_converged": ens["lanczos_converged"], "tuner": ens["tuner"],
                   "parallelism": "1 GPU"},
        "plain_fp64_value": ens.get("plain_fp64_value"), "plain_sample": ens.get("plain_sample"),
        "speedup_vs_plain": ens.get("speedup_vs_plain"),
        "kernel_ms_avg": ens["kernel_ms_avg"],
        "roofline": ens["roofline"],
        "cpu_baseline": cpu,
        "semidilute": semi,
        "charge_hopping": hop,
    }
    print(json.dumps(out, default=float))


if __name__ == "__main__":
    main()
