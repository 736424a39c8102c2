"""Pinning of the oracle's twisted-clover restatement (clover_term.hpp/.cpp,
dirac_operator_tmclovD_eoprec*, dirac_operator_tmclovQ.cpp, cg_invert_tmclovD_eoprec.cpp)
against what the mathematics fixes:

* the clover term of a constant abelian field in each of the six planes equals the
  Sheikholeslami-Wohlert form c_SW (i/4) sigma_munu F_munu = -(c_SW/4) sum_{mu<nu}
  gamma_mu gamma_nu P_munu with P = (Q - Q^+)/4 computed analytically and nissa's
  gamma matrices (dirac.cpp) -- this fixes the leaf orientation, the plane index
  order and the A..F decomposition;
* hermitian chirality blocks and gauge covariance Cl[U^W](x) = W(x) Cl[U](x) W(x)^+
  for a non-abelian field (fixes the order of the products in each leaf);
* at U = 1 (no field strength) every twisted-clover operator is the twisted-mass one;
* D_ee^-1 D_ee = 1 with nissa's block/dagger convention, K(-mu) = K(mu)^+;
* the eo Schur complement of the dense lx operator, and the eo solve solving the
  full lattice system.

The reference's test/projects/reno/comparison_out/2pts (cSW 0.1 on L4T8conf) was
examined as a known-answer test and is not used: its input file is written for
another version of the RIMOM program (keys csw/Walltime/NgaugeConf where
projects/reno_const/RIMOM.cpp:139-204 reads cSW/WallTime/NGaugeConf/UseCGM/
TimeBoundaryCond), and no cSW (nor time boundary) reproduces its P5P5 values
(the first time slice after the source differs by 5e-4 at every cSW).
"""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import G5, GMU, cplx, real, rel_l2, tm_conf, random_su3_field
from oracle import oracle as orc


def _coords(L):
    return np.array(np.meshgrid(*[np.arange(n) for n in L], indexing="ij")).reshape(4, -1).T


def _cdot(a, b):
    return np.vdot(cplx(a).ravel(), cplx(b).ravel())


def _clover_matrix(g, conf, cl, site):
    """12x12 chromo part at one site through the oracle's lx operator:
    g5 (apply_tmclovQ - apply_tmQ) on unit vectors."""
    M = np.zeros((12, 12), complex)
    for k in range(12):
        v = np.zeros((g.V, 4, 3, 2))
        v[site].reshape(12, 2)[k, 0] = 1
        d = g.apply_tmclovQ(conf, 0.2, cl, 0.0, v) - g.apply_tmQ(conf, 0.2, 0.0, v)
        M[:, k] = (G5 @ cplx(d[site]).reshape(4, 3)).ravel()
    return M


@pytest.mark.parametrize("csw", [1.3, -0.7])
def test_clover_term_constant_abelian_fields(csw):
    L = (4, 4, 6, 4)
    g = orc.Geom(L)
    X = _coords(L)
    planes = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for mu, nu in planes:
        # U_nu(x) = diag(e^{i phi x_mu}, e^{-i phi x_mu}, 1), all other links 1: every
        # plaquette (and leaf) of the (mu,nu) plane is W = diag(e^{i phi}, e^{-i phi}, 1)
        phi = 2 * np.pi / L[mu]
        conf = np.zeros((g.V, 4, 3, 3), complex)
        conf[:] = np.eye(3)
        ph = np.exp(1j * phi * X[:, mu])
        conf[:, nu, 0, 0], conf[:, nu, 1, 1] = ph, ph.conj()
        conf = real(conf)
        cl = g.clover_term(conf, csw)
        P = np.diag([2j * np.sin(phi), -2j * np.sin(phi), 0])  # (4W - 4W^+)/4
        expect = -(csw / 4) * np.kron(GMU[mu] @ GMU[nu], P)
        for site in (0, 37, g.V - 1):
            M = _clover_matrix(g, conf, cl, site)
            assert np.abs(M - expect).max() < 1e-14, (mu, nu, site)


def test_clover_blocks_hermitian_and_gauge_covariant():
    L = (4, 4, 2, 6)
    g = orc.Geom(L)
    conf = tm_conf(L, 21)
    csw = 1.1
    cl = cplx(g.clover_term(conf, csw))  # [V][4][3][3]
    for b in (0, 2):  # G and I are hermitian, so each 6x6 chirality block is
        assert np.abs(cl[:, b] - np.conj(np.swapaxes(cl[:, b], -1, -2))).max() < 1e-15
    assert np.abs(cl).max() > 0.05
    om = random_su3_field(L, 22)
    cl2 = cplx(g.clover_term(g.gauge_transform(conf, om), csw))
    W = cplx(om)
    rot = np.einsum("xab,xkbc,xdc->xkad", W, cl, np.conj(W))
    assert np.abs(cl2 - rot).max() < 1e-13


def _free_conf(L, theta_t):
    conf = np.zeros((int(np.prod(L)), 4, 3, 3, 2))
    for c in range(3):
        conf[:, :, c, c, 0] = 1.0
    nq.gauge_phase_lx(L, conf, [theta_t, 0, 0, 0])
    return conf


def test_tmclov_is_tm_without_field_strength():
    L = (4, 2, 4, 2)
    g = orc.Geom(L)
    conf = _free_conf(L, 1.0)
    kappa, mu, csw = 0.14, 0.21, 1.7
    cl = g.clover_term(conf, csw)
    assert np.abs(cl).max() < 1e-15  # rounding of the boundary phase only
    cle, clo = g.split(cl)
    ce, co = g.split(conf)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(g.Vh, 4, 3, 2))
    for dag in (False, True):
        m = -mu if dag else mu
        assert rel_l2(g.tmclovDee(clo, kappa, mu, x, dag=dag), g.tmDee(kappa, m, x)) < 1e-15
        inv = g.tmclov_invert(cle, kappa, mu)
        assert rel_l2(g.inv_tmclovDee(inv, x, dag=dag), g.tmDee(kappa, m, x, inverse=True)) < 1e-14
        assert rel_l2(g.tmclovDkern(ce, co, kappa, clo, inv, mu, x, dag=dag), g.tmDkern(ce, co, kappa, m, x)) < 1e-14
    psi = rng.normal(size=(g.V, 4, 3, 2))
    assert rel_l2(g.apply_tmclovQ(conf, kappa, cl, mu, psi), g.apply_tmQ(conf, kappa, mu, psi)) < 1e-15


def test_tmclov_inverse_and_gamma5_hermiticity():
    L = (4, 4, 2, 4)
    g = orc.Geom(L)
    conf = tm_conf(L, 31)
    kappa, mu, csw = 0.16, 0.08, 1.9
    cl = g.clover_term(conf, csw)
    cle, clo = g.split(cl)
    inv = g.tmclov_invert(cle, kappa, mu)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(g.Vh, 4, 3, 2))
    y = rng.normal(size=(g.Vh, 4, 3, 2))
    for dag in (False, True):
        assert rel_l2(g.inv_tmclovDee(inv, g.tmclovDee(cle, kappa, mu, x, dag=dag), dag=dag), x) < 1e-14
    # Dee(mu)^+ = Dee(-mu): <y, Dee x> = <Dee(dag) y, x>
    lhs = _cdot(y, g.tmclovDee(clo, kappa, mu, x))
    rhs = _cdot(g.tmclovDee(clo, kappa, mu, y, dag=True), x)
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)
    ce, co = g.split(conf)
    k = g.tmclovDkern(ce, co, kappa, clo, inv, mu, x)
    kd = g.tmclovDkern(ce, co, kappa, clo, inv, mu, y, dag=True)
    lhs, rhs = _cdot(y, k), _cdot(kd, x)
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)
    kk = g.tmclovDkern_square(ce, co, kappa, clo, inv, mu, x)
    assert rel_l2(kk, g.tmclovDkern(ce, co, kappa, clo, inv, mu, g.tmclovDkern(ce, co, kappa, clo, inv, mu, x, dag=True))) < 1e-14
    assert _cdot(x, kk).real > 0


def _dense(apply, n_in, shape):
    cols = []
    for kk in range(n_in):
        v = np.zeros(n_in * 2)
        v[2 * kk] = 1
        cols.append(cplx(apply(v.reshape(shape))).ravel())
    return np.array(cols).T


def test_tmclov_eo_schur_complement_dense():
    """K = g5 (D_oo - D_oe D_ee^-1 D_eo) from the dense lx operator D = g5 Q_clov."""
    L = (4, 2, 2, 2)
    g = orc.Geom(L)
    conf = tm_conf(L, 41)
    kappa, mu, csw = 0.16, 0.07, 1.4
    cl = g.clover_term(conf, csw)
    V = g.V
    Q = _dense(lambda v: g.apply_tmclovQ(conf, kappa, cl, mu, v), V * 12, (V, 4, 3, 2))
    D = np.kron(np.eye(V), np.kron(G5, np.eye(3))) @ Q
    par = _coords(L).sum(axis=1) % 2
    idx = lambda p: np.concatenate([np.arange(12) + 12 * s for s in range(V) if par[s] == p])
    ie, io = idx(0), idx(1)
    S = D[np.ix_(io, io)] - D[np.ix_(io, ie)] @ np.linalg.solve(D[np.ix_(ie, ie)], D[np.ix_(ie, io)])
    Kexp = np.kron(np.eye(V // 2), np.kron(G5, np.eye(3))) @ S
    ce, co = g.split(conf)
    cle, clo = g.split(cl)
    inv = g.tmclov_invert(cle, kappa, mu)
    K = _dense(lambda v: g.tmclovDkern(ce, co, kappa, clo, inv, mu, v), g.Vh * 12, (g.Vh, 4, 3, 2))
    assert np.abs(K - Kexp).max() / np.abs(Kexp).max() < 1e-13


def test_tmclov_eo_solve_solves_full_system():
    L = (4, 4, 2, 2)
    g = orc.Geom(L)
    conf = tm_conf(L, 43)
    src = nq.source_lx(L, 9, "gauss")
    sol, st = g.tmclov_solve_eoprec(conf, 0.16, 1.2, -0.04, 1000, 1e-28, src)
    assert st["rc"] == 0 and st["check"] < 1e-26
    # and against the dense solve of the lx system
    cl = g.clover_term(conf, 1.2)
    Q = _dense(lambda v: g.apply_tmclovQ(conf, 0.16, cl, -0.04, v), g.V * 12, (g.V, 4, 3, 2))
    D = np.kron(np.eye(g.V), np.kron(G5, np.eye(3))) @ Q
    xd = np.linalg.solve(D, cplx(src).ravel())
    assert rel_l2(cplx(sol).ravel(), xd) < 1e-12
