"""ctypes binding of libsmdlt_hip.so (the C ABI in include/smdlt_hip.h).

Host-side mirror of the reference's hot-path calls for Python callers (tests,
bench).  The library must have been built (`make -C smdlt-ca-chrg-flow_amd`);
loading fails loudly otherwise -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsmdlt_hip.so")
HEADER = os.path.join(_HERE, "..", "include", "smdlt_hip.h")

FLOW_EQUILIBRIUM, FLOW_SHEAR, FLOW_EXTENSION = 0, 1, 2
PREC_MIXED, PREC_FP64 = 0, 1
NOISE_LANCZOS, NOISE_TEA, NOISE_FREE_DRAINING = 0, 1, 2
MC_REFERENCE, MC_MH_UNIFORM, MC_MH_LEARNED = 0, 1, 2

ERRORS = {0: "ok", -1: "SMDLT_E_ARG", -2: "SMDLT_E_HIP", -3: "SMDLT_E_FENE_OVERSTRETCH",
          -4: "SMDLT_E_LANCZOS_NOCONV", -5: "SMDLT_E_NLIST_OVERFLOW", -6: "SMDLT_E_GEOMETRY",
          -7: "SMDLT_E_STATE", -8: "SMDLT_E_NOMEM", -9: "SMDLT_E_MC_STREAM", -10: "SMDLT_E_NONFINITE"}


class Params(C.Structure):
    _fields_ = [("nb", C.c_int), ("nc", C.c_int), ("L", C.c_double), ("flow_type", C.c_int),
                ("edot", C.c_double), ("dt", C.c_double), ("spring_fene", C.c_int),
                ("epsilon", C.c_double), ("kappab", C.c_double), ("kappas", C.c_double),
                ("qmax", C.c_double), ("lambda_b", C.c_double), ("lambda_d", C.c_double),
                ("barrier", C.c_double), ("barrier2", C.c_double), ("lj_ext_form", C.c_int),
                ("pbc_compat", C.c_int), ("precision", C.c_int), ("noise_mode", C.c_int),
                ("ewald_alpha", C.c_double), ("rc_real", C.c_double), ("skin", C.c_double),
                ("grid", C.c_int * 3), ("spline_order", C.c_int), ("mobility_tol", C.c_double),
                ("lanczos_tol", C.c_double), ("lanczos_maxit", C.c_int), ("seed", C.c_uint64),
                ("hop_radius", C.c_double), ("feature_radius", C.c_double), ("proposal_floor", C.c_double)]


class SmdltError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({lib().smdlt_strerror(code).decode()})")


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C smdlt-ca-chrg-flow_amd`")
        L = C.CDLL(LIB_PATH)
        L.smdlt_version.restype = C.c_char_p
        L.smdlt_strerror.restype = C.c_char_p
        L.smdlt_strerror.argtypes = [C.c_int]
        L.smdlt_default_params.argtypes = [C.POINTER(Params)]
        L.smdlt_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.smdlt_destroy.argtypes = [C.c_void_p]
        L.smdlt_num_beads.argtypes = [C.c_void_p, C.c_int]
        L.smdlt_get_mobility_config.argtypes = [C.c_void_p, C.c_int, _dp]
        L.smdlt_upload_positions.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.smdlt_download_positions.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.smdlt_upload_charges.argtypes = [C.c_void_p, _ip, _ip, _ip]
        L.smdlt_download_charges.argtypes = [C.c_void_p, _ip, _ip]
        L.smdlt_update_lattice.argtypes = [C.c_void_p, C.c_uint64, _dp]
        L.smdlt_remap.argtypes = [C.c_void_p, C.c_uint64]
        L.smdlt_apply_pbc.argtypes = [C.c_void_p]
        L.smdlt_rotate.argtypes = [C.c_void_p, C.c_int]
        L.smdlt_min_image.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp]
        L.smdlt_bin_index.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, _ip, _ip, _ip]
        L.smdlt_neighbors.argtypes = [C.c_void_p, C.c_int, _ip]
        L.smdlt_get_verlet.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_int]
        L.smdlt_forces.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.smdlt_mobility_apply.argtypes = [C.c_void_p, _dp, _dp]
        L.smdlt_brownian.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _ip, _dp]
        L.smdlt_tea_apply.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.smdlt_step.argtypes = [C.c_void_p, C.c_uint64, _dp]
        L.smdlt_run.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int]
        L.smdlt_last_iters.argtypes = [C.c_void_p, _ip]
        L.smdlt_lanczos_status.argtypes = [C.c_void_p, _ip, _ip, _dp]
        L.smdlt_set_tea.argtypes = [C.c_void_p, _dp, _dp]
        L.smdlt_tea_sample.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.smdlt_far_table_dims.argtypes = [C.c_void_p, C.c_int, C.c_double, _ip, _dp]
        L.smdlt_far_table.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp]
        L.smdlt_write_far_table.argtypes = [C.c_char_p, _dp, C.c_int, _ip]
        L.smdlt_stress.argtypes = [C.c_void_p, _dp, C.c_int]
        L.smdlt_chain_observables.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.smdlt_charge_displacements.argtypes = [C.c_void_p, _dp, C.c_int]
        L.smdlt_ran1_seed.argtypes = [C.c_void_p, C.c_int64]
        L.smdlt_ran1_seed.restype = None
        L.smdlt_ran1_fill.argtypes = [C.c_void_p, C.c_int, _fp]
        L.smdlt_ran1_fill.restype = None
        L.smdlt_gasdev_fill.argtypes = [C.c_void_p, C.c_int, _dp]
        L.smdlt_gasdev_fill.restype = None
        L.smdlt_init_chains.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, _dp, _dp, _dp]
        L.smdlt_init_charges.argtypes = [C.c_int, C.c_int, C.c_void_p, _ip]
        L.smdlt_read_restart.argtypes = [C.c_char_p, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_uint64)]
        L.smdlt_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.smdlt_stage_times.argtypes = [C.c_void_p, _dp, _ip, C.c_int]
        L.smdlt_stage_name.restype = C.c_char_p
        L.smdlt_stage_name.argtypes = [C.c_int]
        L.smdlt_sync.argtypes = [C.c_void_p]
        L.smdlt_select_replicas.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.smdlt_mc_sweep.argtypes = [C.c_void_p, C.c_int, _fp, C.c_int, _ip, _ip]
        L.smdlt_mc_sweep_noalloc.argtypes = [C.c_void_p, C.c_int, _fp, C.c_int, _ip, _ip]
        L.smdlt_mc_energy.argtypes = [C.c_void_p, _dp]
        L.smdlt_mlp_set_weights.argtypes = [C.c_void_p, _fp]
        L.smdlt_mlp_loss_grad.argtypes = [C.c_void_p, _fp, C.c_void_p, C.c_void_p, C.c_int, _fp, _fp]
        L.smdlt_mc_features.argtypes = [C.c_void_p, C.c_int, _fp, _ip, _ip, _ip, C.c_int, _ip]
        L.smdlt_mc_record_features.argtypes = [C.c_void_p, C.c_int]
        L.smdlt_tea_avg_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_void_p)]
        L.smdlt_tea_avg_destroy.argtypes = [C.c_void_p]
        L.smdlt_tea_avg_destroy.restype = None
        L.smdlt_tea_avg_add.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_double, _dp, C.c_char_p]
        L.smdlt_tea_avg_write.argtypes = [C.c_void_p, C.c_uint64, C.c_char_p]
        L.smdlt_tea_avg_finish.argtypes = [C.c_void_p, _dp, _dp, _ip]
        L.smdlt_tea_read.argtypes = [C.c_char_p, C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_uint64)]
        L.smdlt_tea_coefficients.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.smdlt_msd.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_char_p, _dp, C.c_int]
        L.smdlt_msd_com.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_int, C.c_char_p, _dp, C.c_int]
        _lib = L
    return _lib


def declared_symbols():
    """Names of every function the ABI header declares."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(smdlt_[a-z0-9_]+)\s*\(", text)))


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _chk(code, where):
    if code < 0:
        raise SmdltError(code, where)
    return code


def default_params(**kw):
    p = Params()
    lib().smdlt_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "grid":
            p.grid = (C.c_int * 3)(*v)
        else:
            setattr(p, k, v)
    return p


class Context:
    """One device context over a list of replicas (smdlt_create/smdlt_destroy)."""

    def __init__(self, params, device=0):
        if isinstance(params, Params):
            params = [params]
        arr = (Params * len(params))(*params)
        h = C.c_void_p()
        _chk(lib().smdlt_create(arr, len(params), device, C.byref(h)), "smdlt_create")
        self.h = h
        self.params = list(params)
        self.nrep = len(params)
        self.n = [p.nb * p.nc for p in params]
        self.ntot = sum(self.n)

    def close(self):
        if self.h:
            lib().smdlt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mobility_config(self, r=0):
        out = np.zeros(9)
        _chk(lib().smdlt_get_mobility_config(self.h, r, _ptr(out, C.c_double)), "config")
        return dict(alpha=float(out[0]), rc=float(out[1]), skin=float(out[2]), grid=tuple(int(x) for x in out[3:6]),
                    order=int(out[6]), list_entries=int(out[7]), cluster_pairs=int(out[8]))

    def upload_positions(self, rx, ry, rz):
        rx, ry, rz = (np.ascontiguousarray(a, np.float64) for a in (rx, ry, rz))
        _chk(lib().smdlt_upload_positions(self.h, _ptr(rx, C.c_double), _ptr(ry, C.c_double),
                                          _ptr(rz, C.c_double)), "upload_positions")

    def download_positions(self):
        out = [np.zeros(self.ntot) for _ in range(3)]
        _chk(lib().smdlt_download_positions(self.h, *(_ptr(a, C.c_double) for a in out)), "download")
        return out

    def update_lattice(self, t):
        out = np.zeros(8 * self.nrep)
        _chk(lib().smdlt_update_lattice(self.h, t, _ptr(out, C.c_double)), "update_lattice")
        return out.reshape(self.nrep, 8)

    def remap(self, t):
        _chk(lib().smdlt_remap(self.h, t), "remap")

    def apply_pbc(self):
        _chk(lib().smdlt_apply_pbc(self.h), "apply_pbc")

    def rotate(self, inverse):
        _chk(lib().smdlt_rotate(self.h, int(inverse)), "rotate")

    def min_image(self, r, i, j):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        out = np.zeros(3 * len(i))
        _chk(lib().smdlt_min_image(self.h, r, len(i), _ptr(i, C.c_int), _ptr(j, C.c_int),
                                   _ptr(out, C.c_double)), "min_image")
        return out.reshape(-1, 3)

    def bin_index(self, r, bin_size, i, j):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        out = np.zeros(len(i), np.int32)
        _chk(lib().smdlt_bin_index(self.h, r, bin_size, len(i), _ptr(i, C.c_int), _ptr(j, C.c_int),
                                   _ptr(out, C.c_int)), "bin_index")
        return out

    def neighbors(self, force=False):
        n = C.c_int()
        _chk(lib().smdlt_neighbors(self.h, int(force), C.byref(n)), "neighbors")
        return n.value

    def verlet(self, r, stride=None):
        n = self.n[r]
        stride = stride or n
        nl = np.zeros(n, np.int32)
        lst = np.zeros(n * stride, np.int32)
        _chk(lib().smdlt_get_verlet(self.h, r, _ptr(nl, C.c_int), _ptr(lst, C.c_int), stride), "get_verlet")
        return nl, lst.reshape(n, stride)

    def forces(self):
        f = [np.zeros(self.ntot) for _ in range(3)]
        _chk(lib().smdlt_forces(self.h, *(_ptr(a, C.c_double) for a in f)), "forces")
        return np.stack(f, 1)

    def mobility_apply(self, v):
        v = np.ascontiguousarray(v, np.float64).reshape(-1)
        u = np.zeros_like(v)
        _chk(lib().smdlt_mobility_apply(self.h, _ptr(v, C.c_double), _ptr(u, C.c_double)), "mobility_apply")
        return u

    def brownian(self, z, tol=1e-6, maxit=30):
        z = np.ascontiguousarray(z, np.float64).reshape(-1)
        y = np.zeros_like(z)
        it = np.zeros(self.nrep, np.int32)
        res = np.zeros(self.nrep)
        code = lib().smdlt_brownian(self.h, _ptr(z, C.c_double), _ptr(y, C.c_double), tol, maxit,
                                    _ptr(it, C.c_int), _ptr(res, C.c_double))
        if code < 0 and code != -4:
            raise SmdltError(code, "brownian")
        return y, it, res, code

    def tea_apply(self, beta, Cv, z):
        beta = np.ascontiguousarray(np.broadcast_to(beta, (self.nrep,)), np.float64)
        Cv = np.ascontiguousarray(Cv, np.float64).reshape(-1)
        z = np.ascontiguousarray(z, np.float64).reshape(-1)
        y = np.zeros_like(z)
        _chk(lib().smdlt_tea_apply(self.h, _ptr(beta, C.c_double), _ptr(Cv, C.c_double), _ptr(z, C.c_double),
                                   _ptr(y, C.c_double)), "tea_apply")
        return y

    def step(self, t, noise=None):
        if noise is None:
            code = lib().smdlt_step(self.h, t, None)
        else:
            noise = np.ascontiguousarray(noise, np.float64).reshape(-1)
            code = lib().smdlt_step(self.h, t, _ptr(noise, C.c_double))
        if code < 0 and code != -4:
            raise SmdltError(code, "step")
        return code

    def run(self, t0, nsteps, use_graph=True):
        code = lib().smdlt_run(self.h, t0, nsteps, int(use_graph))
        if code < 0 and code != -4:
            raise SmdltError(code, "run")
        return code

    def last_iters(self):
        it = np.zeros(self.nrep, np.int32)
        _chk(lib().smdlt_last_iters(self.h, _ptr(it, C.c_int)), "last_iters")
        return it

    def tea_sample(self):
        """smdlt_tea_sample: (beta[nrep], C[3N], eps[nrep]) from the current positions and lattice."""
        beta = np.zeros(self.nrep)
        eps = np.zeros(self.nrep)
        Cv = np.zeros(3 * self.ntot)
        _chk(lib().smdlt_tea_sample(self.h, _ptr(beta, C.c_double), _ptr(Cv, C.c_double), _ptr(eps, C.c_double)),
             "tea_sample")
        return beta, Cv, eps

    def far_table(self, replica=0, nt_avg=8, bin_size=2.0):
        """smdlt_far_table: (table[nt_avg, nbins, 3, 3], dims, bin sizes) of gridHI."""
        dims = np.zeros(3, np.int32)
        sizes = np.zeros(3)
        _chk(lib().smdlt_far_table_dims(self.h, replica, bin_size, _ptr(dims, C.c_int), _ptr(sizes, C.c_double)),
             "far_table_dims")
        nb = int(np.prod(dims))
        table = np.zeros(nt_avg * 9 * nb)
        _chk(lib().smdlt_far_table(self.h, replica, nt_avg, bin_size, _ptr(table, C.c_double)), "far_table")
        return table.reshape(nt_avg, nb, 3, 3), tuple(int(d) for d in dims), sizes

    def stress(self, reset=False):
        """smdlt_stress: per replica (xx, yy, xy) virial sums of the steps since the last reset."""
        tau = np.zeros(3 * self.nrep)
        _chk(lib().smdlt_stress(self.h, _ptr(tau, C.c_double), int(bool(reset))), "stress")
        return tau.reshape(self.nrep, 3)

    def chain_observables(self):
        """smdlt_chain_observables: (ext[nchains], com[nchains, 3], rg[nchains, 6]) of all chains."""
        nk = sum(p.nc for p in self.params)
        ext, com, rg = np.zeros(nk), np.zeros(3 * nk), np.zeros(6 * nk)
        _chk(lib().smdlt_chain_observables(self.h, _ptr(ext, C.c_double), _ptr(com, C.c_double),
                                           _ptr(rg, C.c_double)), "chain_observables")
        return ext, com.reshape(nk, 3), rg.reshape(nk, 6)

    def charge_displacements(self, reset=False):
        """smdlt_charge_displacements: per charge (dx, dy, dz) of the log since the last reset."""
        tot = int(self.nq.sum()) if hasattr(self, "nq") else 0
        d = np.zeros(3 * max(tot, 1))
        _chk(lib().smdlt_charge_displacements(self.h, _ptr(d, C.c_double), int(bool(reset))), "charge_displacements")
        return d[:3 * tot].reshape(-1, 3)

    def set_tea(self, beta, Cv):
        beta = np.ascontiguousarray(np.broadcast_to(beta, (self.nrep,)), np.float64)
        Cv = np.ascontiguousarray(Cv, np.float64).reshape(-1)
        _chk(lib().smdlt_set_tea(self.h, _ptr(beta, C.c_double), _ptr(Cv, C.c_double)), "set_tea")

    def lanczos_status(self):
        it = np.zeros(self.nrep, np.int32)
        stt = np.zeros(self.nrep, np.int32)
        ch = np.zeros(self.nrep)
        _chk(lib().smdlt_lanczos_status(self.h, _ptr(it, C.c_int), _ptr(stt, C.c_int), _ptr(ch, C.c_double)),
             "lanczos_status")
        return it, stt, ch

    def set_profiling(self, on):
        _chk(lib().smdlt_set_profiling(self.h, int(on)), "set_profiling")

    def stage_times(self):
        ms = np.zeros(16)
        cnt = np.zeros(16, np.int32)
        n = _chk(lib().smdlt_stage_times(self.h, _ptr(ms, C.c_double), _ptr(cnt, C.c_int), 16), "stage_times")
        return {lib().smdlt_stage_name(i).decode(): (ms[i], int(cnt[i])) for i in range(n)}

    def select_replicas(self, b, e):
        _chk(lib().smdlt_select_replicas(self.h, b, e), "select_replicas")

    def sync(self):
        _chk(lib().smdlt_sync(self.h), "sync")

    # ---- charges / MC
    def upload_charges(self, nq, track, sign=None):
        nq = np.ascontiguousarray(nq, np.int32)
        track = np.ascontiguousarray(track, np.int32)
        sign = np.ones_like(track) if sign is None else np.ascontiguousarray(sign, np.int32)
        _chk(lib().smdlt_upload_charges(self.h, _ptr(nq, C.c_int), _ptr(track, C.c_int), _ptr(sign, C.c_int)),
             "upload_charges")
        self.nq = nq.copy()

    def download_charges(self):
        tot = int(self.nq.sum())
        tr = np.zeros(max(tot, 1), np.int32)
        old = np.zeros(max(tot, 1), np.int32)
        _chk(lib().smdlt_download_charges(self.h, _ptr(tr, C.c_int), _ptr(old, C.c_int)), "download_charges")
        return tr[:tot], old[:tot]

    def mc_sweep(self, mode=MC_REFERENCE, uniforms=None, noalloc=False):
        """chargeHop for every replica; noalloc=True calls smdlt_mc_sweep_noalloc."""
        tot = int(self.nq.sum())
        consumed = np.zeros(self.nrep, np.int32)
        log = np.zeros(6 * max(tot, 1), np.int32)
        fn = lib().smdlt_mc_sweep_noalloc if noalloc else lib().smdlt_mc_sweep
        if uniforms is None:
            code = fn(self.h, mode, None, 0, _ptr(consumed, C.c_int), _ptr(log, C.c_int))
        else:
            u = np.ascontiguousarray(uniforms, np.float32)
            nu = u.shape[-1]
            code = fn(self.h, mode, _ptr(u, C.c_float), nu, _ptr(consumed, C.c_int), _ptr(log, C.c_int))
        _chk(code, "mc_sweep_noalloc" if noalloc else "mc_sweep")
        return consumed, log[:6 * tot].reshape(-1, 6)

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _chk(lib().smdlt_mlp_set_weights(self.h, _ptr(w, C.c_float)), "mlp_set_weights")

    def mlp_loss_grad(self, w, X=None, y=None, n=None):
        """mean BCE loss and gradient; X = y = None reuses the device-resident batch of size n."""
        w = np.ascontiguousarray(w, np.float32)
        g = np.zeros(len(w), np.float32)
        loss = np.zeros(1, np.float32)
        if X is None:
            code = lib().smdlt_mlp_loss_grad(self.h, _ptr(w, C.c_float), None, None, int(n), _ptr(loss, C.c_float),
                                             _ptr(g, C.c_float))
        else:
            X = np.ascontiguousarray(X, np.float32).reshape(-1, 16)
            y = np.ascontiguousarray(y, np.float32).reshape(-1)
            code = lib().smdlt_mlp_loss_grad(self.h, _ptr(w, C.c_float), X.ctypes.data_as(C.c_void_p),
                                             y.ctypes.data_as(C.c_void_p), len(y), _ptr(loss, C.c_float),
                                             _ptr(g, C.c_float))
        _chk(code, "mlp_loss_grad")
        return float(loss[0]), g

    def mlp_train(self, X, y, w0, iters=600, lr=3e-3):
        """Adam on the GPU loss gradient (the batch is uploaded once); returns (w, first loss, last loss)."""
        w = np.asarray(w0, np.float64).copy()
        m = np.zeros_like(w)
        v = np.zeros_like(w)
        l0, g = self.mlp_loss_grad(w, X, y)
        loss = l0
        for it in range(1, iters + 1):
            if it > 1:
                loss, g = self.mlp_loss_grad(w, n=len(y))
            g = g.astype(np.float64)
            m = 0.9 * m + 0.1 * g
            v = 0.999 * v + 0.001 * g * g
            w -= lr * (m / (1 - 0.9 ** it)) / (np.sqrt(v / (1 - 0.999 ** it)) + 1e-8)
        return w.astype(np.float32), l0, loss

    def mc_features(self, replica=0):
        """features [nq][16], source, destination, accepted of the last MH sweep's moves."""
        cap = int(self.nq[replica])
        f = np.zeros((max(cap, 1), 16), np.float32)
        fr, to, acc = (np.zeros(max(cap, 1), np.int32) for _ in range(3))
        n = np.zeros(1, np.int32)
        _chk(lib().smdlt_mc_features(self.h, replica, _ptr(f, C.c_float), _ptr(fr, C.c_int), _ptr(to, C.c_int),
                                     _ptr(acc, C.c_int), cap, _ptr(n, C.c_int)), "mc_features")
        k = int(n[0])
        return f[:k], fr[:k], to[:k], acc[:k]

    def mc_record_features(self, on=True):
        """smdlt_mc_record_features: MH sweeps record their moves' features (default on)."""
        _chk(lib().smdlt_mc_record_features(self.h, 1 if on else 0), "mc_record_features")

    def mc_energy(self):
        e = np.zeros(self.nrep)
        _chk(lib().smdlt_mc_energy(self.h, _ptr(e, C.c_double)), "mc_energy")
        return e


class TeaAverages:
    """CATEA's running averages for one replica (smdlt_tea_avg_*; S/main26.c:1317-1381):
    per strain bin floor(strain/strain_interval), capped at sbins, the sums of beta and of C
    by index within a chain (i mod 3Nb) and the sample count; printTEA's decomp/DC text
    (S/main26.c:1807-1826), resetAverage's means (S/main26.c:1827-1839) and the
    coefficients updateHI reads at a strain (S/main26.c:757-764)."""

    def __init__(self, nb, nc, sbins=20, strain_interval=1.0):
        self.nb, self.nc, self.sbins, self.strain_interval = nb, nc, sbins, strain_interval
        h = C.c_void_p()
        _chk(lib().smdlt_tea_avg_create(nb, nc, sbins, strain_interval, C.byref(h)), "tea_avg_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().smdlt_tea_avg_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def add(self, strain, beta, Cv, t=0, beta_path=None):
        """one sample (smdlt_tea_sample's beta and C[3N] of this replica); returns its bin;
        beta_path gets CATEA's decomp/B record"""
        Cv = np.ascontiguousarray(Cv, np.float64).reshape(-1)
        return _chk(lib().smdlt_tea_avg_add(self.h, int(t), float(strain), float(beta), _ptr(Cv, C.c_double),
                                            beta_path.encode() if beta_path else None), "tea_avg_add")

    def write_decomp(self, path, t):
        _chk(lib().smdlt_tea_avg_write(self.h, int(t), path.encode()), "tea_avg_write")

    def finish(self):
        """resetAverage: per-bin (beta[sbins+1], C[sbins+1, 3Nb], counts); the sums restart"""
        beta = np.zeros(self.sbins + 1)
        Cb = np.zeros((self.sbins + 1) * 3 * self.nb)
        cnt = np.zeros(self.sbins + 1, np.int32)
        _chk(lib().smdlt_tea_avg_finish(self.h, _ptr(beta, C.c_double), _ptr(Cb, C.c_double), _ptr(cnt, C.c_int)),
             "tea_avg_finish")
        return beta, Cb.reshape(self.sbins + 1, 3 * self.nb), cnt


def tea_read(path, nb, sbins=20):
    """The decomp/DC reader (smdlt_tea_read): (beta[sbins+1], C[sbins+1, 3Nb], t)."""
    beta = np.zeros(sbins + 1)
    Cb = np.zeros((sbins + 1) * 3 * nb)
    t = C.c_uint64()
    _chk(lib().smdlt_tea_read(path.encode(), nb, sbins, _ptr(beta, C.c_double), _ptr(Cb, C.c_double), C.byref(t)),
         "tea_read")
    return beta, Cb.reshape(sbins + 1, 3 * nb), t.value


def tea_coefficients(nb, nc, sbins, strain_interval, strain, beta_bins, C_bins):
    """updateHI's (beta, C[3N]) at a strain (smdlt_tea_coefficients) and its bin."""
    bb = np.ascontiguousarray(beta_bins, np.float64).reshape(-1)
    cb = np.ascontiguousarray(C_bins, np.float64).reshape(-1)
    beta = np.zeros(1)
    Cv = np.zeros(3 * nb * nc)
    b = _chk(lib().smdlt_tea_coefficients(nb, nc, sbins, strain_interval, strain, _ptr(bb, C.c_double),
                                          _ptr(cb, C.c_double), _ptr(beta, C.c_double), _ptr(Cv, C.c_double)),
             "tea_coefficients")
    return float(beta[0]), Cv, b


def msd(ch_path, ncharges, tsteps, msd_start, dt, printprops, out_path=None):
    """calcMSD (smdlt_msd): rows (tau, MSD, MSDx, MSDy, MSDz) from a calcCH file."""
    cap = max(1, tsteps)
    out = np.zeros(5 * cap)
    n = _chk(lib().smdlt_msd(ch_path.encode(), ncharges, tsteps, msd_start, dt, printprops,
                             out_path.encode() if out_path else None, _ptr(out, C.c_double), cap), "msd")
    return out[:5 * n].reshape(-1, 5)


def msd_com(com_path, tsteps, dt, printprops, out_path=None):
    """calcMSDcom (smdlt_msd_com): rows (tau, MSD, MSDx, MSDy, MSDz) from "t dx dy dz" lines."""
    cap = max(1, tsteps)
    out = np.zeros(5 * cap)
    n = _chk(lib().smdlt_msd_com(com_path.encode(), tsteps, dt, printprops, out_path.encode() if out_path else None,
                                 _ptr(out, C.c_double), cap), "msd_com")
    return out[:5 * n].reshape(-1, 5)


# ------------------------------------------------------------- output files
# The reference's text outputs, appended per output step as its writers do
# (C %lf is six decimals).  Callers choose the cadence (printperiod, printprops).

def append_visc(path, t, tau, printperiod, flow_type):
    """printVisc: shear "t tauxy/printperiod" (S/main26.c:1866-1880), planar
    extension "t tauxx/printperiod tauyy/printperiod" (X/main14.c:1886-1897);
    tau is one replica's (xx, yy, xy) row of Context.stress(reset=True)."""
    with open(path, "a") as f:
        if flow_type == FLOW_EXTENSION:
            f.write(f"{t} {tau[0] / printperiod:.6f} {tau[1] / printperiod:.6f}\n")
        else:
            f.write(f"{t} {tau[2] / printperiod:.6f}\n")


def append_ch(path, t, old, track, disp):
    """calcCH (S/main26.c:2452-2490): per charge "t j Charge_old Charge_track dx dy dz"
    with the window sums of Context.charge_displacements (zeros at t = 0)."""
    with open(path, "a") as f:
        for j, (o, k, d) in enumerate(zip(old, track, disp)):
            f.write(f"{t} {j} {o} {k} {d[0]:.6f} {d[1]:.6f} {d[2]:.6f}\n")


def append_ext(path, t, strain, ext, nb, qmax):
    """printExt (S/main26.c:963-974): per chain "t strain x/L (x/L)^2" with
    x/L = ext / (nb qmax), ext from Context.chain_observables."""
    with open(path, "a") as f:
        for e in ext:
            x = e / (nb * qmax)
            f.write(f"{t} {strain:.6f} {x:.6f} {x ** 2:.6f}\n")


def append_rg(path, t, rg):
    """calcRg's file (S/main26.c:942-962): per chain "t i xx xy xz yy yz zz"."""
    with open(path, "a") as f:
        for i, g in enumerate(rg):
            f.write(f"{t} {i} " + " ".join(f"{v:.6f}" for v in g) + "\n")


def append_xyz_frame(path, t, lattice_row, L, rx, ry, rz, charge):
    """printTrajectory (S/main26.c:870-901): extended-xyz frame with the lattice
    L1 = (L1x, L1y), L2 = (L2x, L2y) (a row of Context.update_lattice) and the box
    origin; charge 0 beads as A, charge 1 beads as B (other values are not
    written, as in the reference)."""
    L1x, L1y, L2x, L2y = lattice_row[:4]
    with open(path, "a") as f:
        f.write(f"{len(rx)}\n")
        f.write(f'Lattice="{L1x:.6f} {L1y:.6f} {0.0:.6f} {L2x:.6f} {L2y:.6f} {0.0:.6f} {0.0:.6f} {0.0:.6f} {L:.6f}" '
                f'Origin="{-(L1x + L2x) / 2:.6f} {-(L1y + L2y) / 2:.6f} {-L / 2:.6f}" Properties="" Time={t}\n')
        for x, y, z, q in zip(rx, ry, rz, charge):
            if q == 0:
                f.write(f"A {x:.6f} {y:.6f} {z:.6f}\n")
            elif q == 1:
                f.write(f"B {x:.6f} {y:.6f} {z:.6f}\n")


def write_restart(path, t, rx, ry, rz):
    """printRestart (S/main26.c:902-924): "N", "t", then "A x y z" per bead."""
    with open(path, "w") as f:
        f.write(f"{len(rx)}\n{t}\n")
        for x, y, z in zip(rx, ry, rz):
            f.write(f"A {x:.6f} {y:.6f} {z:.6f}\n")


class Ran1State(C.Structure):
    _fields_ = [("idum", C.c_int64), ("idum2", C.c_int64), ("iy", C.c_int64), ("iv", C.c_int64 * 32),
                ("iset", C.c_int), ("gset", C.c_float)]


class RefStream:
    """The reference's ran1/gasdev stream and the initial state it draws
    (smdlt_ran1_seed/_fill, smdlt_gasdev_fill, smdlt_init_chains, smdlt_init_charges)."""

    def __init__(self, seed=830062920):
        self.state = Ran1State()
        lib().smdlt_ran1_seed(C.byref(self.state), seed)

    def uniforms(self, n):
        out = np.zeros(n, np.float32)
        lib().smdlt_ran1_fill(C.byref(self.state), n, _ptr(out, C.c_float))
        return out

    def gauss(self, n):
        out = np.zeros(n)
        lib().smdlt_gasdev_fill(C.byref(self.state), n, _ptr(out, C.c_double))
        return out

    def init_chains(self, nb, nc, L):
        r = [np.zeros(nb * nc) for _ in range(3)]
        _chk(lib().smdlt_init_chains(nb, nc, L, C.byref(self.state), *(_ptr(a, C.c_double) for a in r)),
             "init_chains")
        return r

    def init_charges(self, n, ncharges):
        track = np.zeros(max(ncharges, 1), np.int32)
        _chk(lib().smdlt_init_charges(n, ncharges, C.byref(self.state), _ptr(track, C.c_int)), "init_charges")
        return track[:ncharges]


def read_restart(path, n):
    """smdlt_read_restart: (rx, ry, rz, t) of the last frame of a printRestart file."""
    r = [np.zeros(n) for _ in range(3)]
    t = C.c_uint64(0)
    _chk(lib().smdlt_read_restart(str(path).encode(), n, *(_ptr(a, C.c_double) for a in r), C.byref(t)),
         "read_restart")
    return r[0], r[1], r[2], int(t.value)


def write_far_table(path, table, dims):
    """smdlt_write_far_table: the reference's mm/G text format."""
    table = np.ascontiguousarray(table, np.float64).reshape(-1)
    nt = table.size // (9 * int(np.prod(dims)))
    d = np.ascontiguousarray(dims, np.int32)
    _chk(lib().smdlt_write_far_table(str(path).encode(), _ptr(table, C.c_double), nt, _ptr(d, C.c_int)),
         "write_far_table")


def mlp_init(seed=0):
    """He-style initial weights of the hop-proposal network, packed W1[64][16], b1,
    W2[64][64], b2, W3[64], b3 (FP32)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / 16), (64, 16))
    W2 = rng.normal(0, np.sqrt(2 / 64), (64, 64))
    W3 = rng.normal(0, np.sqrt(1 / 64), 64)
    return np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), W3, [0.0]]).astype(np.float32)


def stage_names():
    return [lib().smdlt_stage_name(i).decode() for i in range(8)]
