"""Synthetic workloads of the benchmark configurations, shared by bench.py and
the tests: seeded self-avoiding chains grown like the reference's initChains
(bond 2.05a, minimum separation between beads), the parameters of each
configuration and the uneven replica ensemble.  Host-side set-up only; the
HIP library does the work.
"""
import numpy as np

import smdlt_abi as S


def grow_chains(specs, seed, bond=2.05, min_sep=2.0, back=5, patience=40):
    """Seeded self-avoiding random walks for several boxes at once, like the
    reference's initChains (S/main26.c:398-470): bond 2.05a, a bead is redrawn
    while it lies within min_sep of any bead already placed in its box (minimum
    image, cubic box), positions wrapped into [-L/2, L/2).  specs: list of
    (nb, nc, L); returns one (nb*nc, 3) array per box.  All chains grow in step
    (one proposal per unfinished chain per round, vectorised); placed beads sit
    in a cell grid of side >= min_sep, so a proposal is checked against 27 cells.
    A chain that fails `patience` times in a row takes back its last `back`
    beads (self-trapping)."""
    rng = np.random.default_rng(seed)
    ch_rep, ch_nb, ch_first, rep_off = [], [], [], []
    off = 0
    for r, (nb, nc, L) in enumerate(specs):
        rep_off.append(off)
        for c in range(nc):
            ch_rep.append(r)
            ch_nb.append(nb)
            ch_first.append(off + c * nb)
        off += nb * nc
    ch_rep, ch_nb, ch_first = map(np.array, (ch_rep, ch_nb, ch_first))
    pos = np.zeros((off, 3))
    Ls = np.array([s[2] for s in specs], float)
    ncell = np.maximum(3, np.floor(Ls / min_sep).astype(np.int64))
    cbase = np.concatenate([[0], np.cumsum(ncell ** 3)])
    CAP = 5  # points >= min_sep apart in a cell of side >= min_sep: at most 4
    occ = np.full((cbase[-1], CAP), -1, np.int32)
    cnt = np.zeros(cbase[-1], np.int64)
    bead_cell = np.full(off, -1, np.int64)
    offs = np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)])
    jn = np.zeros(len(ch_nb), np.int64)
    fails = np.zeros(len(ch_nb), np.int64)
    while True:
        act = np.nonzero(jn < ch_nb)[0]
        if not act.size:
            break
        rep = ch_rep[act]
        L = Ls[rep][:, None]
        j = jn[act]
        v = rng.normal(size=(act.size, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        prev = pos[ch_first[act] + np.maximum(j - 1, 0)]
        p = np.where((j == 0)[:, None], (rng.random((act.size, 3)) - 0.5) * L, prev + bond * v)
        p -= L * np.round(p / L)
        n = ncell[rep]
        f = np.floor((p / L + 0.5) * n[:, None]).astype(np.int64) % n[:, None]
        nbc = (f[:, None, :] + offs[None]) % n[:, None, None]
        flat = cbase[rep][:, None] + (nbc[..., 0] * n[:, None] + nbc[..., 1]) * n[:, None] + nbc[..., 2]
        cand = occ[flat].reshape(act.size, -1)
        valid = cand >= 0
        d = pos[np.where(valid, cand, 0)] - p[:, None, :]
        d -= L[:, None] * np.round(d / L[:, None])
        ok = ~(valid & (np.einsum("ijk,ijk->ij", d, d) < min_sep * min_sep)).any(axis=1)
        # proposals of one box against each other (act is grouped by box)
        pos_in = np.arange(act.size) - np.searchsorted(rep, rep, side="left")
        G = int(pos_in.max()) + 1
        tab = np.full((len(specs), G), -1)
        tab[rep, pos_in] = np.arange(act.size)
        mate = tab[rep]
        mv = (mate >= 0) & (np.arange(G)[None, :] < pos_in[:, None])
        dd = p[np.where(mv, mate, 0)] - p[:, None, :]
        dd -= L[:, None] * np.round(dd / L[:, None])
        ok &= ~(mv & (np.einsum("ijk,ijk->ij", dd, dd) < min_sep * min_sep)).any(axis=1)
        qs = np.nonzero(ok)[0]
        b = ch_first[act[qs]] + j[qs]
        c = flat[qs, 13]
        pos[b] = p[qs]
        bead_cell[b] = c
        o = np.argsort(c, kind="stable")
        cs = c[o]
        slot = cnt[cs] + np.arange(len(cs)) - np.searchsorted(cs, cs, side="left")
        if slot.size and slot.max() >= CAP:
            raise RuntimeError("grow_chains: cell capacity exceeded")
        occ[cs, slot] = b[o]
        np.add.at(cnt, cs, 1)
        jn[act[ok]] += 1
        fails[act[ok]] = 0
        bad = act[~ok]
        fails[bad] += 1
        for g in bad[fails[bad] > patience]:
            k = min(back, jn[g])
            for bb in range(ch_first[g] + jn[g] - k, ch_first[g] + jn[g]):
                cc = bead_cell[bb]
                row = occ[cc]
                i = int(np.nonzero(row[:cnt[cc]] == bb)[0][0])
                row[i] = row[cnt[cc] - 1]
                row[cnt[cc] - 1] = -1
                cnt[cc] -= 1
                bead_cell[bb] = -1
            jn[g] -= k
            fails[g] = 0
    return [pos[o:o + s[0] * s[1]] for o, s in zip(rep_off, specs)]


def _grow_group(args):
    specs, seed = args
    return grow_chains(specs, seed)


def grow_chains_grouped(specs, seed, groups=32, workers=None):
    """grow_chains over a large ensemble: the boxes are split into `groups`
    contiguous groups of about equal bead count, group g grown with seed
    (seed, g), in worker processes (spawned, so a parent that already holds the
    GPU is never forked).  The result depends on seed and groups only, not on
    the number of workers."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import os

    n = np.array([s[0] * s[1] for s in specs], float)
    cut = np.searchsorted(np.cumsum(n), np.linspace(0, n.sum(), groups + 1)[1:-1])
    bounds = [0] + sorted(set(int(c) + 1 for c in cut if 0 <= c < len(specs) - 1)) + [len(specs)]
    jobs = [(specs[a:b], [seed, g]) for g, (a, b) in enumerate(zip(bounds[:-1], bounds[1:])) if b > a]
    workers = workers or max(1, min(16, os.cpu_count() or 1, len(jobs)))
    if workers == 1:
        parts = [_grow_group(j) for j in jobs]
    else:
        with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            parts = list(ex.map(_grow_group, jobs))
    return [p for part in parts for p in part]


def chains(nb, nc, L, seed):
    return grow_chains([(nb, nc, L)], seed)[0]


def c2_params(prec):
    return S.default_params(nb=200, nc=40, L=80.0, flow_type=S.FLOW_SHEAR, edot=50.0 / 175.0, dt=5e-4,
                            spring_fene=0, kappas=200.0, qmax=2.0, epsilon=0.1, lambda_b=1.0, lambda_d=5.0,
                            precision=prec, lanczos_tol=1e-6, lanczos_maxit=128, seed=0x5EED0000)


def c4_specs(nrep, seed=11):
    """Uneven replicas: chain length U[50, 1000], chain count U{1..20}, box at
    volume fraction 0.02 (L = (N 4 pi / 3 / 0.02)^(1/3)), shear or planar
    extension with probability 1/2, Wi U[1, 100] (edot = Wi / 175)."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrep):
        nb = int(rng.integers(50, 1001))
        nc = int(rng.integers(1, 21))
        L = (nb * nc * 4.0 * np.pi / 3.0 / 0.02) ** (1.0 / 3.0)
        flow = S.FLOW_SHEAR if rng.random() < 0.5 else S.FLOW_EXTENSION
        wi = rng.uniform(1.0, 100.0)
        out.append(dict(nb=nb, nc=nc, L=L, flow=flow, edot=wi / 175.0))
    return out


def to_lab_frame(ctx, specs, pos):
    """Chains are grown in a cube with cubic minimum image.  At t = 0 the
    Kraynik-Reinelt lattice is that cube rotated by theta0 (the reference's
    31.7 degrees), so a planar-extension replica's conformation is the grown one
    rotated into the lab frame; tiled with the cubic lattice instead it would
    overlap its own images across the rotated faces."""
    lat = ctx.update_lattice(0)
    out = []
    for r, sp in enumerate(specs):
        p = pos[r]
        if sp["flow"] == S.FLOW_EXTENSION:
            l1x, l1y = lat[r, 0], lat[r, 1]
            ct, st = l1x / np.hypot(l1x, l1y), l1y / np.hypot(l1x, l1y)
            p = np.stack([p[:, 0] * ct - p[:, 1] * st, p[:, 0] * st + p[:, 1] * ct, p[:, 2]], 1)
        out.append(p)
    return out


def c4_params(specs, prec):
    return [S.default_params(nb=sp["nb"], nc=sp["nc"], L=sp["L"], flow_type=sp["flow"], edot=sp["edot"], dt=5e-4,
                             spring_fene=0, kappas=200.0, qmax=2.0, epsilon=0.1, precision=prec, lanczos_tol=1e-6,
                             lanczos_maxit=128, seed=0x5EED0000 + r) for r, sp in enumerate(specs)]


def c2_charges(n, seed=7):
    rng = np.random.default_rng(seed)
    track = rng.choice(n, n // 2, replace=False).astype(np.int32)
    sign = np.where(np.arange(len(track)) < len(track) // 2, 1, -1).astype(np.int32)
    return track, sign
