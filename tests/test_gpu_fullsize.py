"""GPU checks at the full sizes of the benchmark workloads, through properties
that do not need a dense oracle: the mixed-precision product against a tight
FP64 evaluation (the 1e-5 bound), symmetry of M, and ||M^1/2 z||^2 = z.M.z for
the Lanczos result.  8000 beads in L = 80 under shear (the semidilute
polyelectrolyte line) and 50,000 beads in L = 160 under planar extension."""
import numpy as np
import pytest
from helpers import S, random_walk

pytestmark = pytest.mark.gpu


def xyz(pos):
    return [np.ascontiguousarray(pos[:, k]) for k in range(3)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def params(nb, nc, L, flow, edot, dt, prec, **kw):
    return S.default_params(nb=nb, nc=nc, L=L, flow_type=flow, edot=edot, dt=dt, spring_fene=0, kappas=200.0,
                            qmax=2.0, epsilon=0.1, precision=prec, lanczos_tol=1e-6, lanczos_maxit=128, **kw)


def pair(nb, nc, L, flow, edot, dt, seed, t, ref_tol):
    pos = random_walk(nb, nc, L, seed)
    mixed = S.Context(params(nb, nc, L, flow, edot, dt, S.PREC_MIXED))
    p64 = params(nb, nc, L, flow, edot, dt, S.PREC_FP64)
    p64.mobility_tol = ref_tol
    tight = S.Context(p64)
    for c in (mixed, tight):
        c.upload_positions(*xyz(pos))
        c.remap(t)  # both in the lattice frame of step t
    return mixed, tight, pos


@pytest.fixture(scope="module")
def c2():
    mixed, tight, pos = pair(200, 40, 80.0, S.FLOW_SHEAR, 50.0 / 175.0, 5e-4, 31, 700, 1e-9)
    yield mixed, tight, pos
    mixed.close()
    tight.close()


def test_semidilute_shear_product_within_1e5_of_tight_fp64(c2):
    mixed, tight, pos = c2
    rng = np.random.default_rng(1)
    for _ in range(2):
        v = rng.normal(size=pos.size)
        assert rel(mixed.mobility_apply(v), tight.mobility_apply(v)) < 1e-5


def test_semidilute_shear_product_is_symmetric(c2):
    mixed, tight, pos = c2
    rng = np.random.default_rng(2)
    v, w = rng.normal(size=pos.size), rng.normal(size=pos.size)
    for c, tol in ((tight, 1e-12), (mixed, 1e-6)):
        a, b = w @ c.mobility_apply(v), v @ c.mobility_apply(w)
        assert abs(a - b) <= tol * np.linalg.norm(v) * np.linalg.norm(w)


def test_semidilute_shear_lanczos_norm_identity(c2):
    # ||M^1/2 z||^2 = z.M.z; the stopping tolerance is 1e-6 on the Krylov coefficients
    mixed, tight, pos = c2
    z = np.random.default_rng(3).normal(size=pos.size)
    for c in (mixed, tight):
        y, it, res, code = c.brownian(z, 1e-6, 128)
        assert code == 0 and it[0] < 128
        assert y @ y == pytest.approx(z @ c.mobility_apply(z), rel=1e-5)


def test_planar_extension_50k_product_within_1e5_of_tight_fp64():
    mixed, tight, pos = pair(250, 200, 160.0, S.FLOW_EXTENSION, 20.0 / 175.0, 2e-4, 32, 1300, 1e-7)
    v = np.random.default_rng(4).normal(size=pos.size)
    assert rel(mixed.mobility_apply(v), tight.mobility_apply(v)) < 1e-5
    mixed.close()
    tight.close()
