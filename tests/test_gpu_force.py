"""GPU parity of the rooted-staggered force (summ_the_rootst_eoimpr_quark_force) against
the oracle through the C ABI, FP64 and mixed (FP32 inner, FP64 reliable updates)
multi-shift solves, and the summing into an existing force."""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import cplx, real, rel_l2
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _inputs(L, seed):
    g = orc.Geom(L)
    conf = nq.gauge_random_lx(L, seed, 0.3)
    rng = np.random.default_rng(seed)
    u1 = np.exp(1j * rng.uniform(-0.3, 0.3, size=(g.V, 4)))
    u1[:, 0] *= np.exp(1j * np.pi / L[0])
    X = np.array(np.meshgrid(*[np.arange(n) for n in L], indexing="ij")).reshape(4, -1).T
    eta = np.ones((g.V, 4))
    for mu in range(1, 4):
        eta[:, mu] = eta[:, mu - 1] * (1 - 2 * (X[:, mu - 1] % 2))
    phased = real(cplx(conf) * (u1 * eta)[:, :, None, None])
    pf = rng.normal(size=(g.Vh, 3, 2))
    return g, conf, u1, phased, pf


@pytest.mark.parametrize("L,prec", [((4, 6, 4, 8), "f64"), ((8, 8, 8, 8), "f64"), ((8, 8, 8, 8), "f32")])
def test_rootst_force_matches_oracle(L, prec):
    g, conf, u1, phased, pf = _inputs(L, 3)
    ce, co = g.split(conf)
    u1e, u1o = g.split(real(u1))
    poles = [0.0004, 0.003, 0.02, 0.25, 1.7]
    weights = [0.01, 0.04, 0.15, 0.5, 2.0]
    Fe_ref, Fo_ref, rc = g.rootst_force(ce, co, u1e, u1o, pf, poles, weights, 20000, 1e-28)
    assert rc == 0
    ctx = nq.Ctx(L)
    gh = nq.Gauge(ctx, lx=phased)
    fpf = ctx.field(nq.COLOR, nq.EVEN, pf)
    rng = np.random.default_rng(9)
    Fe0 = rng.normal(size=Fe_ref.shape)
    Fo0 = rng.normal(size=Fo_ref.shape)
    Fe, Fo, st = nq.st_rootst_force(ctx, gh, u1e, u1o, poles, weights, 20000, 1e-26, fpf, Fe0.copy(), Fo0.copy(),
                                    prec=prec)
    assert st["rc"] == nq.NQ_OK
    tol = 1e-10
    assert rel_l2(Fe - Fe0, Fe_ref) < tol and rel_l2(Fo - Fo0, Fo_ref) < tol
    ctx.close()
