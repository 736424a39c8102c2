"""Pinning of the oracle's rooted-staggered force (summ_the_rootst_eoimpr_quark_force,
hmc/fermions/rootst_eoimpr_quark_force.cpp:17-85) by what the mathematics fixes: it is
the gradient of the pseudofermion action S(U) = sum_k w_k <pf, (p_k - 1/4 Deo Doe)^-1 pf>
built on the links with background field and staggered phases.  For a link moved
as U -> exp(i eps H) U (H hermitian) the central difference of S must equal
Re tr(i H U F) for every link, parity, direction and generator checked.
"""
import numpy as np

import nissa_amd as nq
from helpers import cplx, real
from oracle import oracle as orc


def _expm_i(H, eps):
    w, V = np.linalg.eigh(H)
    return (V * np.exp(1j * eps * w)) @ V.conj().T


def test_rootst_force_is_the_action_gradient():
    L = (4, 4, 2, 4)
    g = orc.Geom(L)
    conf = nq.gauge_random_lx(L, 7, 0.3)  # links without phases
    ce, co = g.split(conf)
    rng = np.random.default_rng(1)
    u1 = np.exp(1j * rng.uniform(-0.3, 0.3, size=(g.V, 4)))  # background U(1) field
    u1[:, 0] *= np.exp(1j * np.pi / L[0])  # antiperiodic time
    u1e, u1o = g.split(real(u1))
    pf = rng.normal(size=(g.Vh, 3, 2))
    poles, weights = [0.01, 0.1, 0.7], [0.3, 0.5, 1.1]
    Fe, Fo, rc = g.rootst_force(ce, co, u1e, u1o, pf, poles, weights, 10000, 1e-30)
    assert rc == 0
    X = np.array(np.meshgrid(*[np.arange(n) for n in L], indexing="ij")).reshape(4, -1).T
    eta = np.ones((g.V, 4))
    for mu in range(1, 4):  # eta_mu(x) = prod_{nu < mu} (-1)^{x_nu}
        eta[:, mu] = eta[:, mu - 1] * (1 - 2 * (X[:, mu - 1] % 2))
    phe, pho = g.split(real(u1 * eta))

    def action(c):
        pe = real(cplx(c[0]) * cplx(phe)[:, :, None, None])
        po = real(cplx(c[1]) * cplx(pho)[:, :, None, None])
        sols, _ = g.st_cgm(pe, po, poles, 10000, [1e-30] * 3, pf)
        return sum(w * np.vdot(cplx(pf).ravel(), cplx(s).ravel()).real for w, s in zip(weights, sols))

    gens = [np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]]), np.array([[0, -1j, 0], [1j, 0, 0], [0, 0, 0]]),
            np.diag([1.0, -1.0, 0.0]), np.array([[0, 0, 0], [0, 0, -1j], [0, 1j, 0]])]
    eps = 1e-5
    for trial in range(8):
        par, i, mu = trial % 2, int(rng.integers(g.Vh)), int(rng.integers(4))
        H = gens[trial % 4]
        S = []
        for sgn in (1, -1):
            c = [ce.copy(), co.copy()]
            c[par][i, mu] = real(_expm_i(H, sgn * eps) @ cplx(c[par][i, mu]))
            S.append(action(c))
        dS = (S[0] - S[1]) / (2 * eps)
        U, F = cplx((ce, co)[par][i, mu]), cplx((Fe, Fo)[par][i, mu])
        pred = np.trace(1j * H @ U @ F).real
        assert abs(dS - pred) < 1e-7 * max(1.0, abs(pred)), (trial, dS, pred)
