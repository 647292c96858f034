"""ctypes view of the CPU oracle (oracle/build/libsmdlt_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.  The product library never
loads it.

Each wrapper names the reference lines its C function restates
(S/ = shear/src/, X/ = extensional/src/).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libsmdlt_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class RNG(C.Structure):
    _fields_ = [("idum", C.c_int64), ("idum2", C.c_int64), ("iy", C.c_int64),
                ("iv", C.c_int64 * 32), ("iset", C.c_int), ("gset", C.c_float)]


class Lattice(C.Structure):
    _fields_ = [("flow", C.c_int), ("L", C.c_double), ("edot", C.c_double), ("dt", C.c_double),
                ("tp", C.c_int), ("point0", (C.c_double * 2) * 3), ("point", (C.c_double * 2) * 3),
                ("L1", C.c_double * 2), ("L2", C.c_double * 2),
                ("L1xp", C.c_double), ("L1yp", C.c_double), ("L2xp", C.c_double), ("L2yp", C.c_double),
                ("detL", C.c_double), ("L2pmag", C.c_double), ("treal", C.c_int),
                ("delx", C.c_double), ("theta", C.c_double), ("ct", C.c_double), ("st", C.c_double),
                ("pbc_compat", C.c_int)]


class Bins(C.Structure):
    _fields_ = [("bin_size_x", C.c_double), ("bin_size_y", C.c_double), ("bin_size_z", C.c_double),
                ("nbx", C.c_int), ("nby", C.c_int), ("nbz", C.c_int), ("nhx", C.c_int),
                ("nhy", C.c_int), ("nhz", C.c_int), ("num_bins", C.c_int)]


class Cell(C.Structure):
    _fields_ = [("a1x", C.c_double), ("a1y", C.c_double), ("a2x", C.c_double),
                ("a2y", C.c_double), ("Lz", C.c_double)]


class MCState(C.Structure):
    _fields_ = [("n", C.c_int), ("nb", C.c_int), ("nc", C.c_int), ("nq", C.c_int),
                ("track", C.POINTER(C.c_int)), ("indices", C.POINTER(C.c_int)),
                ("old", C.POINTER(C.c_int)), ("ivalues", C.POINTER(C.c_int)),
                ("charge", C.POINTER(C.c_int))]


class RefRun(C.Structure):
    _fields_ = [("nb", C.c_int), ("nc", C.c_int), ("flow", C.c_int),
                ("L", C.c_double), ("edot", C.c_double), ("dt", C.c_double), ("kappas", C.c_double),
                ("qmax", C.c_double), ("epsilon", C.c_double), ("lambda_b", C.c_double),
                ("lambda_d", C.c_double), ("barrier", C.c_double), ("barrier2", C.c_double),
                ("spring_fene", C.c_int), ("nq", C.c_int), ("mstep", C.c_int), ("srpy", C.c_int),
                ("nt_avg", C.c_int), ("beta", C.c_double), ("seed", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_rng_seed.argtypes = [C.POINTER(RNG), C.c_int64]
        L.orc_bf16_round.restype = C.c_float
        L.orc_bf16_round.argtypes = [C.c_float]
        L.orc_mh_features.argtypes = [C.POINTER(Lattice), C.POINTER(MHParams), _dp, _dp, _dp, C.c_void_p, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_mlp_forward.restype = C.c_double
        L.orc_mlp_forward.argtypes = [C.c_void_p, _dp]
        L.orc_mh_sweep.argtypes = [C.POINTER(Lattice), C.POINTER(MHParams), _dp, _dp, _dp, C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, _fp, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_ran1.argtypes = [C.POINTER(RNG)]
        L.orc_ran1.restype = C.c_float
        L.orc_gasdev.argtypes = [C.POINTER(RNG)]
        L.orc_gasdev.restype = C.c_float
        L.orc_ran1_fill.argtypes = [C.POINTER(RNG), C.c_int, _fp]
        L.orc_gasdev_fill.argtypes = [C.POINTER(RNG), C.c_int, _dp]
        L.orc_init_chains.argtypes = [C.POINTER(RNG), C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
        L.orc_init_chains.restype = C.c_int
        L.orc_init_charges.argtypes = [C.POINTER(RNG), C.c_int, C.c_int, _ip]
        L.orc_init_charges.restype = None
        L.orc_lattice_init.argtypes = [C.POINTER(Lattice), C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_lattice_update.argtypes = [C.POINTER(Lattice), C.c_uint64]
        L.orc_apply_pbc.argtypes = [C.POINTER(Lattice), C.c_int, _dp, _dp, _dp]
        L.orc_rotate.argtypes = [C.POINTER(Lattice), C.c_int, _dp, _dp, C.c_int]
        L.orc_nid.argtypes = [C.POINTER(Lattice), _dp, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_nid_bin.argtypes = [C.POINTER(Lattice), _dp, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_bins_init.argtypes = [C.POINTER(Bins), C.POINTER(Lattice), C.c_double]
        L.orc_binij.argtypes = [C.POINTER(Bins), C.POINTER(Lattice), _dp, _dp, _dp, C.c_int, C.c_int]
        L.orc_binij.restype = C.c_int
        L.orc_verlet.argtypes = [C.POINTER(Lattice), C.c_int, _dp, _dp, _dp, C.c_double, _ip, _ip, C.c_int]
        L.orc_check_verlet.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double]
        L.orc_check_verlet.restype = C.c_int
        L.orc_bond_force.argtypes = [C.POINTER(Lattice), C.c_int, C.c_int, _dp, _dp, _dp, C.c_int,
                                     C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_bond_force.restype = C.c_int
        L.orc_chain_obs.argtypes = [C.POINTER(Lattice), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_chain_obs.restype = None
        L.orc_lj_force.argtypes = [C.POINTER(Lattice), C.c_int, _dp, _dp, _dp, _ip, _ip, C.c_int,
                                   C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_coulomb_force.argtypes = [C.POINTER(Lattice), C.c_int, _dp, _dp, _dp, _ip, C.c_double,
                                        C.c_double, _dp, _dp, _dp]
        L.orc_coulomb_energy.argtypes = [C.POINTER(Lattice), C.c_int, _dp, _dp, _dp, _ip, C.c_double, C.c_double]
        L.orc_coulomb_energy.restype = C.c_double
        L.orc_rpy_real_coeffs.argtypes = [C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_ewald_bin.argtypes = [C.POINTER(Lattice), C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), _dp, _dp, C.POINTER(C.c_int)]
        L.orc_catea.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.orc_catea.restype = C.c_double
        L.orc_far_table.argtypes = [C.POINTER(Lattice), C.c_int, C.POINTER(Bins), C.c_double, _dp]
        L.orc_rpy_ewald_dense.argtypes = [C.POINTER(Cell), C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_rpy_ewald_matvec.argtypes = [C.POINTER(Cell), C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                           _dp, _dp]
        L.orc_rpy_free_dense.argtypes = [C.c_int, _dp, _dp]
        L.orc_ewald_selfconst.argtypes = [C.c_double]
        L.orc_ewald_selfconst.restype = C.c_double
        L.orc_calchi_ref.argtypes = [C.POINTER(Lattice), C.c_int, _dp, _dp, _dp, C.c_uint64, C.c_int,
                                     C.POINTER(Bins), C.c_void_p, C.c_double, _dp,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_tea_apply.argtypes = [C.c_int, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_sym_sqrt_apply.argtypes = [C.c_int, _dp, _dp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_sym_sqrt_apply.restype = C.c_int
        L.orc_lanczos_sqrt.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, C.POINTER(C.c_int)]
        L.orc_lanczos_sqrt.restype = C.c_int
        L.orc_matvec.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_charge_hop.argtypes = [C.POINTER(MCState), C.POINTER(Lattice), _dp, _dp, _dp, C.c_double,
                                     C.c_double, C.c_double, C.c_double, _fp, C.c_int,
                                     C.POINTER(C.c_int), C.c_void_p, C.POINTER(C.c_double)]
        L.orc_charge_hop.restype = C.c_int
        L.orc_ref_run.argtypes = [C.POINTER(RefRun), _dp, _dp, _dp, C.c_int, C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]
        L.orc_ref_run.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


# ---------------------------------------------------------------- RNG
class Ran1:
    """ran1/gasdev stream (S/Initialization.c:601-634, S/main26.c:515-540)."""

    def __init__(self, seed):
        self.s = RNG()
        lib().orc_rng_seed(C.byref(self.s), int(seed))

    def copy(self):
        """An independent stream in the same state (ran1's statics and gasdev's
        cached deviate included)."""
        other = Ran1.__new__(Ran1)
        other.s = RNG()
        C.memmove(C.byref(other.s), C.byref(self.s), C.sizeof(RNG))
        return other

    def uniforms(self, n):
        out = np.empty(n, np.float32)
        lib().orc_ran1_fill(C.byref(self.s), n, out)
        return out

    def gauss(self, n):
        out = np.empty(n, np.float64)
        lib().orc_gasdev_fill(C.byref(self.s), n, out)
        return out

    def init_chains(self, nb, nc, L):
        """initChains for a new run (S/main26.c:346-470)."""
        r = [np.zeros(nb * nc) for _ in range(3)]
        assert lib().orc_init_chains(C.byref(self.s), nb, nc, L, *r) == 0
        return r

    def init_charges(self, n, ncharges):
        """the charge placement of initChains (S/main26.c:485-500)."""
        track = np.zeros(max(ncharges, 1), np.int32)
        lib().orc_init_charges(C.byref(self.s), n, ncharges, track)
        return track[:ncharges]


# ---------------------------------------------------------------- lattice
def lattice(flow, L, edot, dt, pbc_compat=1, t=0):
    """initLattice + updateLattice(t) (S/Initialization.c:217-298, S/main26.c:1382-1408,
    X/main14.c:1426-1450)."""
    lat = Lattice()
    lib().orc_lattice_init(C.byref(lat), flow, L, edot, dt, pbc_compat)
    if t:
        lib().orc_lattice_update(C.byref(lat), t)
    return lat


def lattice_update(lat, t):
    lib().orc_lattice_update(C.byref(lat), int(t))


def apply_pbc(lat, rx, ry, rz):
    rx, ry, rz = (np.ascontiguousarray(a, np.float64).copy() for a in (rx, ry, rz))
    lib().orc_apply_pbc(C.byref(lat), len(rx), rx, ry, rz)
    return rx, ry, rz


def rotate(lat, rx, ry, inverse):
    rx, ry = (np.ascontiguousarray(a, np.float64).copy() for a in (rx, ry))
    lib().orc_rotate(C.byref(lat), len(rx), rx, ry, int(inverse))
    return rx, ry


def nid(lat, rx, ry, rz, i, j):
    out = np.zeros(3)
    lib().orc_nid(C.byref(lat), rx, ry, rz, i, j, out)
    return out


def nid_bin(lat, rx, ry, rz, i, j):
    out = np.zeros(3)
    lib().orc_nid_bin(C.byref(lat), rx, ry, rz, i, j, out)
    return out


def bins(lat, bin_size):
    b = Bins()
    lib().orc_bins_init(C.byref(b), C.byref(lat), bin_size)
    return b


def binij(b, lat, rx, ry, rz, i, j):
    return lib().orc_binij(C.byref(b), C.byref(lat), rx, ry, rz, i, j)


def verlet(lat, rx, ry, rz, rv, stride=None):
    n = len(rx)
    stride = stride or n
    nl = np.zeros(n, np.int32)
    lst = np.zeros(n * stride, np.int32)
    lib().orc_verlet(C.byref(lat), n, rx, ry, rz, rv, nl, lst, stride)
    return nl, lst.reshape(n, stride)


# ---------------------------------------------------------------- forces
def bond_force(lat, nb, nc, rx, ry, rz, fene, kappas, qmax, kappab=0.0):
    n = nb * nc
    f = [np.zeros(n) for _ in range(3)]
    tau = np.zeros(3)
    st = lib().orc_bond_force(C.byref(lat), nb, nc, rx, ry, rz, int(fene), kappas, qmax, kappab,
                              f[0], f[1], f[2], tau)
    return st, np.stack(f, 1), tau


def lj_force(lat, rx, ry, rz, nlist, lst, r2cut, epsilon, ext_form):
    n = len(rx)
    lst = np.ascontiguousarray(lst, np.int32)
    stride = lst.shape[1]
    f = [np.zeros(n) for _ in range(3)]
    tau = np.zeros(3)
    lib().orc_lj_force(C.byref(lat), n, rx, ry, rz, np.ascontiguousarray(nlist, np.int32),
                       lst.reshape(-1), stride, r2cut, epsilon, int(ext_form), f[0], f[1], f[2], tau)
    return np.stack(f, 1), tau


def chain_obs(lat, nb, nc, rx, ry, rz):
    """Per-chain extension, centre of mass and gyration tensor of the unwrapped
    chains (calcExt/calcCOM/calcRg, S/main26.c:74-93,925-962,2391-2415)."""
    ext, com, rg = np.zeros(nc), np.zeros(3 * nc), np.zeros(6 * nc)
    lib().orc_chain_obs(C.byref(lat), nb, nc, *(np.ascontiguousarray(a, np.float64) for a in (rx, ry, rz)),
                        ext, com, rg)
    return ext, com.reshape(nc, 3), rg.reshape(nc, 6)


def coulomb_force(lat, rx, ry, rz, q, lambda_b, lambda_d):
    n = len(rx)
    f = [np.zeros(n) for _ in range(3)]
    lib().orc_coulomb_force(C.byref(lat), n, rx, ry, rz, np.ascontiguousarray(q, np.int32),
                            lambda_b, lambda_d, f[0], f[1], f[2])
    return np.stack(f, 1)


def coulomb_energy(lat, rx, ry, rz, q, lambda_b, lambda_d):
    return lib().orc_coulomb_energy(C.byref(lat), len(rx), rx, ry, rz,
                                    np.ascontiguousarray(q, np.int32), lambda_b, lambda_d)


# ---------------------------------------------------------------- mobility
def rpy_real_coeffs(r, alpha, overlap_mode=0):
    t1, t2 = C.c_double(), C.c_double()
    lib().orc_rpy_real_coeffs(r, alpha, overlap_mode, C.byref(t1), C.byref(t2))
    return t1.value, t2.value


def cell_of(lat):
    """lattice vectors in the frame positions are wrapped in: a1=(L1xp, 0), a2=(L2xp, L2yp)."""
    return Cell(lat.L1xp, 0.0, lat.L2xp, lat.L2yp, lat.L)


def ewald_dense(cell, pos, alpha, rcut=None, kcut=None, overlap_mode=0):
    """Exact Ewald-summed RPY matrix (3n x 3n).  Defaults put the truncation
    error near FP64 round-off."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
    n = len(pos) // 3
    rcut = rcut if rcut is not None else 6.2 / alpha
    kcut = kcut if kcut is not None else 12.4 * alpha
    M = np.zeros((3 * n) * (3 * n))
    lib().orc_rpy_ewald_dense(C.byref(cell), n, pos, alpha, rcut, kcut, overlap_mode, M)
    return M.reshape(3 * n, 3 * n)


def ewald_dense_alpha(cell):
    """Splitting parameter that minimises the per-pair work of ewald_dense
    (real-space images within 6.2/alpha plus wave vectors within 12.4 alpha)."""
    a1 = np.hypot(cell.a1x, cell.a1y)
    a2 = np.hypot(cell.a2x, cell.a2y)
    area = abs(cell.a1x * cell.a2y - cell.a1y * cell.a2x)
    V = area * cell.Lz
    w = np.array([area / a2, area / a1, cell.Lz])
    best = None
    for alpha in np.geomspace(0.03, 3.0, 80):
        rcut, kcut = 6.2 / alpha, 12.4 * alpha
        images = np.prod(2 * (np.ceil(rcut / w) + 1) + 1)
        nk = 4.19 * kcut ** 3 * V / (2 * np.pi) ** 3
        if best is None or images + nk < best[0]:
            best = (images + nk, alpha)
    return best[1]


def ewald_matvec_alpha(cell, n):
    """Splitting parameter that minimises a rough cost model of ewald_matvec
    (pair distance tests and real-space terms against structure-factor terms)."""
    a1 = np.hypot(cell.a1x, cell.a1y)
    a2 = np.hypot(cell.a2x, cell.a2y)
    area = abs(cell.a1x * cell.a2y - cell.a1y * cell.a2x)
    V = area * cell.Lz
    w = np.array([area / max(a1, a2), area / min(a1, a2), cell.Lz])
    best = None
    for alpha in np.geomspace(0.03, 3.0, 80):
        rcut, kcut = 6.2 / alpha, 12.4 * alpha
        nimg = np.prod(2 * np.floor(rcut / w + 0.5) + 1)
        real = n * n * nimg * 2e-9 + n * min(n * nimg, 4.19 * rcut ** 3 * n / V) * 6e-8
        nk = 4.19 * kcut ** 3 * V / (2 * np.pi) ** 3 / 2
        cost = real + 2 * n * nk * 4e-9
        if best is None or cost < best[0]:
            best = (cost, alpha)
    return best[1]


def ewald_matvec(cell, pos, v, alpha=None, rcut=None, kcut=None, overlap_mode=0):
    """u = M v for the matrix of ewald_dense, computed without forming M."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
    n = len(pos) // 3
    alpha = alpha if alpha is not None else ewald_matvec_alpha(cell, n)
    rcut = rcut if rcut is not None else 6.2 / alpha
    kcut = kcut if kcut is not None else 12.4 * alpha
    u = np.zeros(3 * n)
    lib().orc_rpy_ewald_matvec(C.byref(cell), n, pos, alpha, rcut, kcut, overlap_mode,
                               np.ascontiguousarray(v, np.float64), u)
    return u


def rpy_free_dense(pos):
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
    n = len(pos) // 3
    M = np.zeros((3 * n) * (3 * n))
    lib().orc_rpy_free_dense(n, pos, M)
    return M.reshape(3 * n, 3 * n)


def selfconst(alpha):
    return lib().orc_ewald_selfconst(alpha)


def calchi_ref(lat, rx, ry, rz, t, nt_avg, b, table, box_volume):
    n = len(rx)
    M = np.zeros(9 * n * n)
    far = C.c_int()
    kc = C.c_int()
    tp = None
    if table is not None:
        table = np.ascontiguousarray(table, np.float64)
        tp = table.ctypes.data_as(C.c_void_p)
    lib().orc_calchi_ref(C.byref(lat), n, rx, ry, rz, int(t), nt_avg, C.byref(b), tp, box_volume, M,
                         C.byref(far), C.byref(kc))
    return M.reshape(3 * n, 3 * n), far.value, kc.value


def ewald_bin(lat, rx, ry, rz, nb, box_volume, m_hi=3.5):
    """ewaldBin (S/main26.c:975-1260): pair sum eps, rowsum[3N], self block Dtr[3], k count.
    m_hi = 3.5 is the reference's split."""
    rx, ry, rz = (np.ascontiguousarray(a, np.float64) for a in (rx, ry, rz))
    n = len(rx)
    eps = C.c_double()
    rowsum = np.zeros(3 * n)
    dtr = np.zeros(3)
    kc = C.c_int()
    lib().orc_ewald_bin(C.byref(lat), n, nb, rx, ry, rz, box_volume, m_hi, C.byref(eps), rowsum, dtr, C.byref(kc))
    return eps.value, rowsum, dtr, kc.value


def far_table(lat, nt_avg, b, box_volume):
    """gridHI (S/gridHI.c:3-244): table[nt_avg, nbins, 3, 3]; lat is the lattice at t = 0."""
    nb = b.nbx * b.nby * b.nbz
    table = np.zeros(nt_avg * 9 * nb)
    lib().orc_far_table(C.byref(lat), nt_avg, C.byref(b), box_volume, table)
    return table.reshape(nt_avg, nb, 3, 3)


def catea(n, eps, rowsum):
    """CATEA (S/main26.c:1317-1381) for one sample: beta and C[3N]."""
    Cv = np.zeros(3 * n)
    beta = lib().orc_catea(n, eps, np.ascontiguousarray(rowsum, np.float64), Cv)
    return beta, Cv


def tea_apply(M, beta, Cv, R):
    n = M.shape[0] // 3
    y = np.zeros(3 * n)
    lib().orc_tea_apply(n, np.ascontiguousarray(M), beta, np.ascontiguousarray(Cv, np.float64),
                        np.ascontiguousarray(R, np.float64), y)
    return y


def sym_sqrt_apply(A, z):
    n3 = A.shape[0]
    y = np.zeros(n3)
    lo, hi = C.c_double(), C.c_double()
    st = lib().orc_sym_sqrt_apply(n3, np.ascontiguousarray(A), np.ascontiguousarray(z, np.float64), y,
                                  C.byref(lo), C.byref(hi))
    return y, st, lo.value, hi.value


def lanczos_sqrt(A, z, tol, maxit):
    n3 = A.shape[0]
    y = np.zeros(n3)
    it = C.c_int()
    st = lib().orc_lanczos_sqrt(n3, np.ascontiguousarray(A), np.ascontiguousarray(z, np.float64), tol,
                                maxit, y, C.byref(it))
    return y, st, it.value


# ---------------------------------------------------------------- MC
class ChargeHop:
    """chargeHop state (S/main26.c:1926-2390): Charge_track, Charge_indices,
    Charge_old, i_values, Charge."""

    def __init__(self, nb, nc, track):
        n = nb * nc
        self.track = np.ascontiguousarray(track, np.int32).copy()
        self.indices = self.track.copy()
        self.old = self.track.copy()
        self.ivalues = np.zeros(n, np.int32)
        self.charge = np.zeros(n, np.int32)
        self.charge[self.track] = 1
        P = C.POINTER(C.c_int)
        self.st = MCState(n, nb, nc, len(self.track), self.track.ctypes.data_as(P),
                          self.indices.ctypes.data_as(P), self.old.ctypes.data_as(P),
                          self.ivalues.ctypes.data_as(P), self.charge.ctypes.data_as(P))

    def sweep(self, lat, rx, ry, rz, lambda_b, lambda_d, barrier, barrier2, u):
        u = np.ascontiguousarray(u, np.float32)
        cur = C.c_int(0)
        log = np.zeros(6 * max(1, len(self.track)), np.int32)
        e0 = C.c_double()
        st = lib().orc_charge_hop(C.byref(self.st), C.byref(lat), rx, ry, rz, lambda_b, lambda_d,
                                  barrier, barrier2, u, len(u), C.byref(cur),
                                  log.ctypes.data_as(C.c_void_p), C.byref(e0))
        return st, cur.value, log.reshape(-1, 6), e0.value


# ---------------------------------------------------------------- MH charge hopping
class MHParams(C.Structure):
    _fields_ = [("n", C.c_int), ("nb", C.c_int), ("hop_radius", C.c_double), ("feature_radius", C.c_double),
                ("floor_eps", C.c_double), ("lambda_b", C.c_double), ("lambda_d", C.c_double),
                ("barrier", C.c_double), ("barrier2", C.c_double)]


def mh_params(n, nb, hop_radius=3.0, feature_radius=4.0, floor_eps=0.05, lambda_b=1.0, lambda_d=5.0,
              barrier=0.0, barrier2=0.0):
    return MHParams(n, nb, hop_radius, feature_radius, floor_eps, lambda_b, lambda_d, barrier, barrier2)


def bf16_round(w):
    w = np.ascontiguousarray(w, np.float32)
    return np.array([lib().orc_bf16_round(float(x)) for x in w], np.float32)


def mh_features(lat, p, rx, ry, rz, qbead, src, dst, qk, clr=-1, set_=-1, ncand=0, anchor=None):
    """anchor: the bead the pair distances are taken relative to (default src)."""
    f = np.zeros(16)
    qb = np.ascontiguousarray(qbead, np.int32)
    lib().orc_mh_features(C.byref(lat), C.byref(p), rx, ry, rz, qb.ctypes.data_as(C.c_void_p), src, dst, qk, clr,
                          set_, ncand, src if anchor is None else anchor, f)
    return f


def mlp_forward(w, f):
    w = np.ascontiguousarray(w, np.float32)
    return lib().orc_mlp_forward(w.ctypes.data_as(C.c_void_p), np.ascontiguousarray(f, np.float64))


def mh_sweep(lat, p, rx, ry, rz, track, sign, qbead, w, u, features=False):
    """Runs len(u)//3 attempts in place on track and qbead; returns (accepted, log, feats)."""
    u = np.ascontiguousarray(u, np.float32)
    na = len(u) // 3
    log = np.zeros(6 * max(1, na), np.int32)
    feats = np.zeros(16 * max(1, na), np.float32) if features else None
    wp = None if w is None else np.ascontiguousarray(w, np.float32)
    acc = lib().orc_mh_sweep(C.byref(lat), C.byref(p), rx, ry, rz, len(track), track.ctypes.data_as(C.c_void_p),
                             np.ascontiguousarray(sign, np.int32).ctypes.data_as(C.c_void_p),
                             qbead.ctypes.data_as(C.c_void_p),
                             None if wp is None else wp.ctypes.data_as(C.c_void_p), u, na,
                             log.ctypes.data_as(C.c_void_p),
                             None if feats is None else feats.ctypes.data_as(C.c_void_p))
    return acc, log.reshape(-1, 6)[:na], (feats.reshape(-1, 16)[:na] if features else None)


def ref_run(params: dict, rx, ry, rz, nsteps, nthreads=0):
    p = RefRun(**params)
    rx, ry, rz = (np.ascontiguousarray(a, np.float64).copy() for a in (rx, ry, rz))
    st = C.c_int()
    thi = C.c_double(0.0)
    el = lib().orc_ref_run(C.byref(p), rx, ry, rz, nsteps, nthreads, C.byref(st), C.byref(thi))
    return el, st.value, (rx, ry, rz), thi.value


def num_threads():
    return lib().orc_num_threads()
