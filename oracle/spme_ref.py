"""numpy restatement of the smooth-PME Ewald-RPY mobility product, FP64.

TEST INFRASTRUCTURE ONLY (a checker).  The reference computes the reciprocal
part as a direct k-sum (S/main26.c:1764-1788) and has no grid; this file states
the grid method (Essmann et al. 1995, cardinal B-splines, a deformed lattice
a1=(a1x,0,0), a2=(a2x,a2y,0), a3=(0,0,Lz)) with the Beenakker reciprocal kernel
(S/main26.c:1622) so that the GPU pipeline can be compared stage by stage, and
it is itself pinned against the dense Ewald sum of oracle.ewald_dense.
"""
import numpy as np


def bspline_weights(x, p):
    """W[..., j] = M_p(x + j), j = 0..p-1 (x in [0,1))."""
    x = np.asarray(x, np.float64)
    A = np.zeros(x.shape + (p,))
    A[..., 0] = 1.0
    for k in range(2, p + 1):
        div = 1.0 / (k - 1)
        B = np.zeros_like(A)
        for j in range(p):
            lo = A[..., j - 1] if j >= 1 else 0.0
            B[..., j] = div * ((x + j) * A[..., j] + (k - x - j) * lo)
        A = B
    return A


def bspline_moduli(K, p):
    """|b(m)|^2 = 1/|sum_{k=0}^{p-2} M_p(k+1) e^{2 pi i m k/K}|^2."""
    Mint = bspline_weights(np.zeros(1), p)[0]  # M_p(j), j = 0..p-1
    m = np.arange(K)
    s = np.zeros(K, complex)
    for k in range(p - 1):
        s += Mint[k + 1] * np.exp(2j * np.pi * m * k / K)
    d = np.abs(s) ** 2
    return np.where(d > 1e-10, 1.0 / np.maximum(d, 1e-300), 0.0)


def real_coeffs(r, alpha):
    """Beenakker real-space t1, t2 with the exact overlap correction (r<2)."""
    r = np.asarray(r, np.float64)
    rr = r * r
    a = alpha
    a2, a3 = a * a, a ** 3
    a5, a7 = a ** 5, a ** 7
    from scipy.special import erfc
    ec = erfc(a * r)
    g = np.exp(-a2 * rr) / np.sqrt(np.pi)
    C1 = 0.75 / r + 0.5 / (r * rr)
    C2 = 4 * a7 * rr * rr + 3 * a3 * rr - 20 * a5 * rr - 4.5 * a + 14 * a3 + a / rr
    C3 = 0.75 / r - 1.5 / (r * rr)
    C4 = -4 * a7 * rr * rr - 3 * a3 * rr + 16 * a5 * rr + 1.5 * a - 2 * a3 - 3 * a / rr
    t1 = C1 * ec + C2 * g
    t2 = C3 * ec + C4 * g
    ov = r < 2.0
    t1 = np.where(ov, t1 + (1 - 9 * r / 32) - C1, t1)
    t2 = np.where(ov, t2 + 3 * r / 32 - C3, t2)
    return t1, t2


def min_image(d, cell):
    """Minimum image for a1=(a1x,0), a2=(a2x,a2y), a3=(0,0,Lz) by brute force over
    in-plane neighbours of the rounded image."""
    a1x, a2x, a2y, Lz = cell
    d = np.array(d, np.float64)
    s2 = np.round(d[..., 1] / a2y)
    d[..., 0] -= s2 * a2x
    d[..., 1] -= s2 * a2y
    d[..., 0] -= np.round(d[..., 0] / a1x) * a1x
    d[..., 2] -= np.round(d[..., 2] / Lz) * Lz
    best = d.copy()
    bn = np.sum(d * d, -1)
    for n1 in (-1, 0, 1):
        for n2 in (-1, 0, 1):
            e = d.copy()
            e[..., 0] += n1 * a1x + n2 * a2x
            e[..., 1] += n2 * a2y
            en = np.sum(e * e, -1)
            m = en < bn
            best[m] = e[m]
            bn = np.where(m, en, bn)
    return best


def mobility_apply(pos, v, cell, alpha, rc, K, p):
    """u = M v by the PME split.  pos (n,3), v (n,3); cell = (a1x, a2x, a2y, Lz)."""
    a1x, a2x, a2y, Lz = cell
    pos = np.asarray(pos, np.float64)
    v = np.asarray(v, np.float64).reshape(-1, 3)
    n = len(pos)
    K0, K1, K2 = K
    V = a1x * a2y * Lz
    # fractional coordinates and weights
    s2 = pos[:, 1] / a2y
    s1 = (pos[:, 0] - a2x * s2) / a1x
    s3 = pos[:, 2] / Lz
    S = np.stack([s1, s2, s3], 1)
    u = S * np.array(K, float)
    fl = np.floor(u)
    fr = u - fl
    base = fl.astype(np.int64)
    W = [bspline_weights(fr[:, d], p) for d in range(3)]
    idx = [(base[:, d:d + 1] - np.arange(p)[None, :]) % K[d] for d in range(3)]
    Q = np.zeros((3, K2, K1, K0))
    for i in range(n):
        w = np.einsum("a,b,c->abc", W[2][i], W[1][i], W[0][i])
        for c in range(3):
            Q[c][np.ix_(idx[2][i], idx[1][i], idx[0][i])] += w * v[i, c]
    Qh = np.fft.fftn(Q, axes=(1, 2, 3))  # forward, e^{-}
    m3 = np.fft.fftfreq(K2, 1.0 / K2)
    m2 = np.fft.fftfreq(K1, 1.0 / K1)
    m1 = np.fft.fftfreq(K0, 1.0 / K0)
    M3, M2, M1 = np.meshgrid(m3, m2, m1, indexing="ij")
    kb = (2 * np.pi / a1x, -2 * np.pi * a2x / (a1x * a2y), 2 * np.pi / a2y, 2 * np.pi / Lz)
    kx = M1 * kb[0]
    ky = M1 * kb[1] + M2 * kb[2]
    kz = M3 * kb[3]
    kk = kx * kx + ky * ky + kz * kz
    kk[0, 0, 0] = 1.0
    q = kk / (4 * alpha * alpha)
    g = (1 - kk / 3) * (1 + q + 2 * q * q) * 6 * np.pi / kk * np.exp(-q) / V
    B = (bspline_moduli(K2, p)[:, None, None] * bspline_moduli(K1, p)[None, :, None]
         * bspline_moduli(K0, p)[None, None, :])
    g = g * B
    g[0, 0, 0] = 0.0
    # Nyquist planes dropped (the GPU keeps a real, symmetric operator this way)
    g[K2 // 2, :, :] = 0.0
    g[:, K1 // 2, :] = 0.0
    g[:, :, K0 // 2] = 0.0
    kv = np.stack([kx, ky, kz])
    dot = (kv * Qh).sum(0) / kk
    Ph = g * (Qh - kv * dot)
    P = np.real(np.fft.ifftn(Ph, axes=(1, 2, 3))) * (K0 * K1 * K2)  # unnormalised inverse
    out = np.zeros((n, 3))
    for i in range(n):
        w = np.einsum("a,b,c->abc", W[2][i], W[1][i], W[0][i])
        for c in range(3):
            out[i, c] = np.sum(w * P[c][np.ix_(idx[2][i], idx[1][i], idx[0][i])])
    # self
    sc = 1 - 6 / np.sqrt(np.pi) * alpha + 40 / (3 * np.sqrt(np.pi)) * alpha ** 3
    out += sc * v
    # real space, minimum image pairs within rc
    for i in range(n):
        d = min_image(pos - pos[i], cell)
        r = np.sqrt((d * d).sum(1))
        m = (r < rc) & (np.arange(n) != i)
        if not m.any():
            continue
        t1, t2 = real_coeffs(r[m], alpha)
        dv = (d[m] * v[m]).sum(1)
        out[i] += (t1[:, None] * v[m] + (t2 * dv / r[m] ** 2)[:, None] * d[m]).sum(0)
    return out.reshape(-1)
