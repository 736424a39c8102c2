"""Pin the CPU oracle before trusting it as the checker of the GPU path.

Every check here is independent of the oracle's own statement of the algorithm:
reference fixtures (the L4T8conf plaquette, tmLQCD propagators solved on that
configuration), closed forms (free plane waves with nissa's gamma matrices
written out as data, the free staggered D^2), exact algebraic identities
(gauge covariance, gamma5-hermiticity, antihermiticity of D_st, the eo Schur
complement of the full-lattice operator) and brute force (dense direct solves).
"""
import math

import numpy as np
import pytest

import lime
import nissa_amd as nq
from helpers import G5, GMU, cplx, real, rel_l2, coords, parity, tm_conf, st_conf, random_su3_field
from oracle import oracle as orc

# ---------------------------------------------------------------- geometry


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (8, 4, 6, 2), (2, 6, 4, 8)])
def test_geometry_closed_forms(L):
    g = orc.Geom(L)
    c = coords(L)
    par = parity(L)
    for lx in range(g.V):
        assert g.eo_of_lx(lx) == lx >> 1
        assert g.parity(lx) == par[lx]
        assert g.lx_of_eo(par[lx], lx >> 1) == lx
        for mu in range(4):
            up = c[lx].copy()
            up[mu] = (up[mu] + 1) % L[mu]
            dw = c[lx].copy()
            dw[mu] = (dw[mu] - 1) % L[mu]
            idx = lambda v: ((v[0] * L[1] + v[1]) * L[2] + v[2]) * L[3] + v[3]
            assert g.neigh(lx, mu, True) == idx(up)
            assert g.neigh(lx, mu, False) == idx(dw)


def test_split_paste_inverse_and_product_helpers():
    L = (4, 6, 2, 4)
    g = orc.Geom(L)
    a = np.random.default_rng(3).normal(size=(g.V, 4, 3, 2))
    e, o = g.split(a)
    e2, o2 = nq.split_lx(L, a)
    assert np.array_equal(e, e2) and np.array_equal(o, o2)
    assert np.array_equal(g.paste(e, o), a)
    assert np.array_equal(nq.paste_lx(L, e, o), a)

# ---------------------------------------------------------------- reference fixtures

L4T8 = (8, 4, 4, 4)


def test_plaquette_of_reference_conf(golden_dir):
    # test/nissa/src/plaquette_computation.cpp:7-20: 0.6350055722288719, tolerance 1e-14
    conf = lime.read_ildg_conf(f"{golden_dir}/L4T8conf", L4T8)
    g = orc.Geom(L4T8)
    assert abs(g.plaquette(conf) - 0.6350055722288719) < 1e-14


@pytest.mark.parametrize("fname,kappa,sc", [
    ("source_1.0003.00.inverted.00", 0.150, 0),
    ("source_1.0003.00.inverted.07", 0.150, 7),
    ("source_2.0003.00.inverted.04", 0.167, 4),
])
def test_tmlqcd_propagators_solve_point_source(golden_dir, fname, kappa, sc):
    """tmLQCD CG solutions on L4T8conf (test/tools/meson_2pts, kappa from the file's
    inverter-info, mu=0, antiperiodic time) must satisfy D psi = point source at
    the origin with spin-colour index = file suffix.  This pins the hop term's
    gamma basis, signs, link placement, the boundary phase and the normalisation."""
    conf = lime.read_ildg_conf(f"{golden_dir}/L4T8conf", L4T8)
    info = lime.text_record(f"{golden_dir}/{fname}", "inverter-info")
    assert f"kappa = {kappa:.6f}" in info
    psi = lime.read_scidac_spincolor(f"{golden_dir}/{fname}", L4T8)
    nq.gauge_phase_lx(L4T8, conf, [1.0, 0, 0, 0])
    g = orc.Geom(L4T8)
    D = g.apply_tmQ(conf, kappa, 0.0, psi)
    D[:, 2:] *= -1  # D = g5 Q
    eta = np.zeros_like(D)
    eta[0].reshape(12, 2)[sc, 0] = 1.0
    r = np.sum((D - eta) ** 2) / np.sum(eta ** 2)
    assert r < 1e-12, r

# ---------------------------------------------------------------- free field closed forms


def _free_conf(L, theta_t):
    V = int(np.prod(L))
    conf = np.zeros((V, 4, 3, 3, 2))
    for c in range(3):
        conf[:, :, c, c, 0] = 1.0
    nq.gauge_phase_lx(L, conf, [theta_t, 0, 0, 0])
    return conf


@pytest.mark.parametrize("n,theta", [((1, 2, 0, 3), 1.0), ((0, 1, 1, 1), 0.0), ((3, 0, 2, 1), 0.5)])
def test_free_plane_wave_tmQ(n, theta):
    L = (8, 4, 6, 4)
    kappa, mu = 0.137, 0.21
    conf = _free_conf(L, theta)
    c = coords(L)
    p = np.array([2 * np.pi * n[m] / L[m] for m in range(4)])
    ph = np.exp(1j * (c @ p))
    u = np.random.default_rng(1).normal(size=(4, 3)) + 1j * np.random.default_rng(2).normal(size=(4, 3))
    psi = ph[:, None, None] * u[None]
    # effective momentum including the boundary phase on t links
    pe = p.copy()
    pe[0] += np.pi * theta / L[0]
    one = np.eye(4)
    hop = sum((one - GMU[m]) * np.exp(1j * pe[m]) + (one + GMU[m]) * np.exp(-1j * pe[m]) for m in range(4))
    Dm = (1 / (2 * kappa)) * one + 1j * mu * G5 - 0.5 * hop
    expect = ph[:, None, None] * ((G5 @ Dm) @ u)[None]
    g = orc.Geom(L)
    out = g.apply_tmQ(conf, kappa, mu, real(psi))
    assert rel_l2(cplx(out), expect) < 1e-14
    # the eo hop on odd sites is the hop of the even part of the plane wave
    e, o = g.split(real(psi))
    ce, co = g.split(conf)
    ho = g.tm_hop(1, ce, co, e)
    exp_lx = ph[:, None, None] * (hop @ u)[None]
    _, exp_o = g.split(real(exp_lx))
    assert rel_l2(ho, exp_o) < 1e-14


def test_free_staggered_D2_closed_form():
    """With links c_mu * eta_mu(x), -1/4 Deo Doe = -M^2 and M^2 psi(x) =
    1/4 sum_mu (c^2 psi(x+2mu) + conj(c)^2 psi(x-2mu) - 2 psi(x)) (the cross terms
    cancel only for correct staggered phases)."""
    L = (8, 4, 6, 4)
    conf = np.zeros((int(np.prod(L)), 4, 3, 3, 2))
    for c in range(3):
        conf[:, :, c, c, 0] = 1.0
    nq.gauge_stagphases_lx(L, conf)
    nq.gauge_phase_lx(L, conf, [-1.0, 0, 0, 0])
    cmu = [np.exp(-1j * np.pi / L[0]), 1, 1, 1]
    g = orc.Geom(L)
    ce, co = g.split(conf)
    rng = np.random.default_rng(5)
    psi_lx = rng.normal(size=(g.V, 3)) + 1j * rng.normal(size=(g.V, 3))
    psi_lx[parity(L) == 1] = 0
    field = psi_lx.reshape(L + (3,))
    m2 = 0.0123
    expect = m2 * field
    for m in range(4):
        expect = expect - 0.25 * (cmu[m] ** 2 * np.roll(field, -2, axis=m) + np.conj(cmu[m]) ** 2 *
                                  np.roll(field, 2, axis=m) - 2 * field)
    e, _ = g.split(real(psi_lx))
    out = g.stD2ee_m2(ce, co, m2, e)
    exp_e, _ = g.split(real(expect.reshape(g.V, 3)))
    assert rel_l2(out, exp_e) < 1e-14

# ---------------------------------------------------------------- identities


def test_gauge_covariance_tm_and_staggered():
    L = (4, 4, 6, 4)
    g = orc.Geom(L)
    conf = tm_conf(L, 11)
    om = random_su3_field(L, 12)
    conf2 = g.gauge_transform(conf, om)
    rng = np.random.default_rng(7)
    psi = rng.normal(size=(g.V, 4, 3, 2))
    psi2 = g.color_rotate(psi, om, 4)
    kappa, mu = 0.16, 0.05
    a = g.apply_tmQ(conf2, kappa, mu, psi2)
    b = g.color_rotate(g.apply_tmQ(conf, kappa, mu, psi), om, 4)
    assert rel_l2(a, b) < 1e-13
    # K on odd sites
    ce, co = g.split(conf)
    ce2, co2 = g.split(conf2)
    _, po = g.split(psi)
    _, po2 = g.split(psi2)
    k2 = g.tmDkern(ce2, co2, kappa, mu, po2)
    k1 = g.tmDkern(ce, co, kappa, mu, po)
    _, k1r = g.split(g.color_rotate(g.paste(np.zeros_like(k1), k1), om, 4))
    assert rel_l2(k2, k1r) < 1e-13
    # staggered
    sconf = st_conf(L, 13)
    sconf2 = g.gauge_transform(sconf, om)
    se, so = g.split(sconf)
    se2, so2 = g.split(sconf2)
    chi = rng.normal(size=(g.V, 3, 2))
    chi[parity(L) == 1] = 0
    chi2 = g.color_rotate(chi, om, 1)
    ce_, _ = g.split(chi)
    ce2_, _ = g.split(chi2)
    s2 = g.stD2ee_m2(se2, so2, 0.3, ce2_)
    s1 = g.stD2ee_m2(se, so, 0.3, ce_)
    s1r, _ = g.split(g.color_rotate(g.paste(s1, np.zeros_like(s1)), om, 1))
    assert rel_l2(s2, s1r) < 1e-13


def _cdot(a, b):
    return np.vdot(cplx(a).ravel(), cplx(b).ravel())


def test_gamma5_hermiticity_of_K():
    """K(mu)^+ = K(-mu), hence K(+mu)K(-mu) is Hermitian positive definite."""
    L = (4, 4, 4, 8)
    g = orc.Geom(L)
    ce, co = g.split(tm_conf(L, 21))
    rng = np.random.default_rng(9)
    x = rng.normal(size=(g.Vh, 4, 3, 2))
    y = rng.normal(size=(g.Vh, 4, 3, 2))
    kappa, mu = 0.16, 0.03
    lhs = _cdot(x, g.tmDkern(ce, co, kappa, mu, y))
    rhs = _cdot(g.tmDkern(ce, co, kappa, -mu, x), y)
    assert abs(lhs - rhs) / abs(lhs) < 1e-13
    kk = _cdot(x, g.tmDkern_square(ce, co, kappa, mu, x))
    assert kk.real > 0 and abs(kk.imag) / kk.real < 1e-13
    # <x, KK^+ x> = |K(-mu) x|^2  (used by the device CG for <p, A p>)
    t = g.tmDkern(ce, co, kappa, -mu, x)
    assert abs(kk.real - np.sum(t ** 2)) / kk.real < 1e-13


def test_staggered_antihermiticity():
    L = (4, 4, 4, 8)
    g = orc.Geom(L)
    ce, co = g.split(st_conf(L, 31))
    rng = np.random.default_rng(2)
    xo = rng.normal(size=(g.Vh, 3, 2))
    ye = rng.normal(size=(g.Vh, 3, 2))
    lhs = _cdot(xo, g.st_2Doe(ce, co, ye))
    rhs = _cdot(g.st_Deo_half(ce, co, xo), ye)
    assert abs(lhs + 4 * rhs) / abs(lhs) < 1e-13


def test_staggered_full_D_schur_and_antihermiticity():
    """apply_stD (dirac_operator_stD.hpp:16-41,254-291) against the pinned pieces:
    D(m) on (x_e, 0) is (m x_e, 1/2 2Doe x_e); D^+ D = -D(-m) D(m) restricted to the
    even sites is stD2ee_m2(m^2) (the operator inv_stD_cg inverts); D - m is
    anti-hermitian on the full lattice."""
    L = (4, 6, 4, 8)
    g = orc.Geom(L)
    ce, co = g.split(st_conf(L, 37))
    rng = np.random.default_rng(5)
    xe, xo = rng.normal(size=(g.Vh, 3, 2)), rng.normal(size=(g.Vh, 3, 2))
    ye, yo = rng.normal(size=(g.Vh, 3, 2)), rng.normal(size=(g.Vh, 3, 2))
    z = np.zeros_like(xe)
    m = 0.3
    de, do = g.apply_stD(ce, co, m, xe, z)
    assert rel_l2(de, m * xe) < 1e-15 and rel_l2(do, 0.5 * g.st_2Doe(ce, co, xe)) < 1e-14
    ee, _ = g.apply_stD(ce, co, -m, de, do)
    assert rel_l2(-ee, g.stD2ee_m2(ce, co, m * m, xe)) < 1e-13
    he, ho = g.apply_stD(ce, co, 0.0, xe, xo)
    ke, ko = g.apply_stD(ce, co, 0.0, ye, yo)
    lhs = _cdot(ye, he) + _cdot(yo, ho)
    rhs = _cdot(ke, xe) + _cdot(ko, xo)
    assert abs(lhs + rhs) / abs(lhs) < 1e-13


def _dense(apply, n_in, shape):
    cols = []
    for k in range(n_in):
        for ri in range(2):
            v = np.zeros(n_in * 2)
            v[2 * k + ri] = 1
            cols.append(cplx(apply(v.reshape(shape))).ravel())
    # columns for real and imaginary unit inputs -> complex-linear matrix
    M = np.array(cols[0::2]).T
    Mi = np.array(cols[1::2]).T
    assert np.allclose(Mi, 1j * M, atol=1e-13)
    return M


def test_eo_schur_complement_dense():
    """K(mu) = g5 (D_oo - D_oe D_ee^-1 D_eo) with D = g5 Q assembled column by column
    from the full-lattice operator (a separate code path of the reference)."""
    L = (4, 4, 2, 2)
    g = orc.Geom(L)
    conf = tm_conf(L, 41)
    kappa, mu = 0.16, 0.07
    V = g.V
    Q = _dense(lambda v: g.apply_tmQ(conf, kappa, mu, v), V * 12, (V, 4, 3, 2))
    g5 = np.kron(np.eye(V), np.kron(G5, np.eye(3)))
    D = g5 @ Q
    par = parity(L)
    idx = lambda p: np.concatenate([np.arange(12) + 12 * s for s in range(V) if par[s] == p])
    # eo index order = lx order within each parity
    ie, io = idx(0), idx(1)
    Dee, Deo, Doe, Doo = D[np.ix_(ie, ie)], D[np.ix_(ie, io)], D[np.ix_(io, ie)], D[np.ix_(io, io)]
    S = Doo - Doe @ np.linalg.solve(Dee, Deo)
    g5o = np.kron(np.eye(V // 2), np.kron(G5, np.eye(3)))
    Kexp = g5o @ S
    ce, co = g.split(conf)
    K = _dense(lambda v: g.tmDkern(ce, co, kappa, mu, v), g.Vh * 12, (g.Vh, 4, 3, 2))
    assert np.abs(K - Kexp).max() / np.abs(Kexp).max() < 1e-13
    # square, and a direct dense solve against the oracle CG
    Km = _dense(lambda v: g.tmDkern(ce, co, kappa, -mu, v), g.Vh * 12, (g.Vh, 4, 3, 2))
    KK = _dense(lambda v: g.tmDkern_square(ce, co, kappa, mu, v), g.Vh * 12, (g.Vh, 4, 3, 2))
    assert np.abs(KK - K @ Km).max() / np.abs(KK).max() < 1e-13
    b = np.random.default_rng(3).normal(size=(g.Vh, 4, 3, 2))
    xd = np.linalg.solve(KK, cplx(b).ravel())
    sol, st = g.tm_cg_kernsq(ce, co, kappa, mu, 1000, 1e-28, b)
    assert st["rc"] == 0
    assert rel_l2(cplx(sol).ravel(), xd) < 1e-11


def test_eo_solve_solves_full_system():
    """inv_tmD_cg_eoprec: D sol = source on the full lattice (dense check at 4x4x2x2)."""
    L = (4, 4, 2, 2)
    g = orc.Geom(L)
    conf = tm_conf(L, 43)
    kappa, mu = 0.16, -0.04
    src = nq.source_lx(L, 9, "gauss")
    sol, st = g.tm_solve_eoprec(conf, kappa, mu, 1000, 1e-28, src)
    assert st["rc"] == 0 and st["check"] < 1e-26
    Q = _dense(lambda v: g.apply_tmQ(conf, kappa, mu, v), g.V * 12, (g.V, 4, 3, 2))
    g5 = np.kron(np.eye(g.V), np.kron(G5, np.eye(3)))
    xd = np.linalg.solve(g5 @ Q, cplx(src).ravel())
    assert rel_l2(cplx(sol).ravel(), xd) < 1e-11


def test_eo_solve_iteration_count_matches_reference_probe():
    """8^4, kappa 0.16, mu 0.01, Z2 source on even sites, residue 1e-24: the reference
    took 70 CG iterations on a gauge field of the same construction (a run of nissa's
    own inv_tmD_cg_eoprec)."""
    L = (8, 8, 8, 8)
    g = orc.Geom(L)
    conf = tm_conf(L, 1)
    src = nq.source_lx(L, 7, "z2", 12, 1)
    sol, st = g.tm_solve_eoprec(conf, 0.16, 0.01, 10000, 1e-24, src)
    assert st["rc"] == 0
    assert 66 <= st["iters"] <= 74
    assert st["check"] < 2e-24


def test_tmQ2_identity_and_legacy_solvers():
    """Q(-mu)Q(+mu) = Q0^2 + mu^2 (Q(mu) = Q0 + i mu), and the legacy CG/CGM on it."""
    L = (4, 4, 4, 4)
    g = orc.Geom(L)
    conf = tm_conf(L, 51)
    kappa, mu = 0.15, 0.2
    x = nq.source_lx(L, 4, "gauss")
    q2 = g.apply_tmQ2(conf, kappa, mu, x)
    q0 = g.apply_tmQ(conf, kappa, 0.0, g.apply_tmQ(conf, kappa, 0.0, x))
    assert rel_l2(q2, q0 + mu * mu * x) < 1e-14
    masses = [0.05, 0.2, 0.6]
    sols, st = g.tmQ2_cgm(conf, kappa, masses, 5000, [1e-24] * 3, x)
    assert st["rc"] == 0
    for m, s, tr in zip(masses, sols, st["true_res_sq"]):
        ref, st1 = g.tmQ2_cg(conf, kappa, m, 5000, 1e-26, x)
        assert rel_l2(s, ref) < 1e-9
        assert tr < 2e-24


def test_staggered_cgm_matches_single_shift_cg():
    L = (4, 4, 4, 4)
    g = orc.Geom(L)
    ce, co = g.split(st_conf(L, 61))
    src = nq.source_lx(L, 8, "gauss", 3, 1)
    se, _ = g.split(src)
    m2 = 0.0025 ** 2
    shifts = [m2 + 10 ** (-4 + 5 * k / 11) for k in range(0, 12, 3)]
    sols, st = g.st_cgm(ce, co, shifts, 20000, [1e-22] * len(shifts), se)
    assert st["rc"] == 0
    for s, x, tr in zip(shifts, sols, st["true_res_sq"]):
        ref, st1 = g.st_cg(ce, co, s, 20000, 1e-26, se)
        assert st1["rc"] == 0
        assert rel_l2(x, ref) < 1e-8
        assert tr < 2e-22


def test_staggered_rational_sum_against_dense_partial_fractions():
    """summ_src_and_all_inv_stD2ee_m2_cgm (cgm_invert_stD2ee_m2_portable.hpp:31-55):
    chi = c0 eta + sum_k w_k (A + p_k)^-1 eta with A the massless -1/4 Deo Doe, here
    against the same sum formed with dense solves on 4x4x2x2.  Distinct weights of both
    signs and a non-zero c0, so a dropped pole, a flipped weight, a shift of the wrong
    sign or c0 on the wrong vector each change the result; and the partial fractions
    of 1/((x+a)(x+b)) give (A+b)(A+a) chi = eta."""
    L = (4, 4, 2, 2)
    g = orc.Geom(L)
    ce, co = g.split(st_conf(L, 63))
    se, _ = g.split(nq.source_lx(L, 12, "gauss", 3, 1))
    A = _dense(lambda v: g.stD2ee_m2(ce, co, 0.0, v), g.Vh * 3, (g.Vh, 3, 2))
    eta = cplx(se).ravel()
    c0, poles, weights = 0.35, [0.004, 0.09, 1.7], [0.8, -0.25, 2.1]
    chi, st = g.st_rat_summ(ce, co, c0, poles, weights, 20000, 1e-28, se)
    assert st["rc"] == 0
    ref = c0 * eta + sum(w * np.linalg.solve(A + p * np.eye(len(eta)), eta) for p, w in zip(poles, weights))
    assert rel_l2(cplx(chi).ravel(), ref) < 1e-11
    a, b = 0.01, 0.2
    chi2, st2 = g.st_rat_summ(ce, co, 0.0, [a, b], [1 / (b - a), -1 / (b - a)], 20000, 1e-28, se)
    assert st2["rc"] == 0
    z = g.stD2ee_m2(ce, co, b, g.stD2ee_m2(ce, co, a, chi2))
    assert rel_l2(z, se) < 1e-11


def test_quad_precision_defect_correction():
    L = (4, 4, 4, 4)
    g = orc.Geom(L)
    ce, co = g.split(tm_conf(L, 71))
    kappa, mu = 0.16, 0.01
    b = np.random.default_rng(4).normal(size=(g.Vh, 4, 3, 2))
    # the __float128 operator agrees with the double one to rounding
    assert rel_l2(g.tmDkern_square_q(ce, co, kappa, mu, b), g.tmDkern_square(ce, co, kappa, mu, b)) < 1e-15
    xq, st = g.tm_cg128_kernsq(ce, co, kappa, mu, 1e-30, b)
    assert st["rc"] == 0 and st["rel_res_sq"] < 1e-30
    xd, std = g.tm_cg_kernsq(ce, co, kappa, mu, 10000, 1e-24, b)
    assert 1e-14 < rel_l2(xd, xq) < 1e-9  # double solve is accurate but not exact


def test_reductions_exact_and_accurate():
    rng = np.random.default_rng(0)
    a = rng.integers(-1000, 1000, size=(1001, 24)).astype(np.float64)
    b = rng.integers(-1000, 1000, size=(1001, 24)).astype(np.float64)
    exact = int(np.sum(a.astype(np.int64) * b.astype(np.int64)))
    assert orc.scalar_prod(a, b, 24) == exact
    x = rng.normal(size=(12345, 6))
    assert abs(orc.scalar_prod(x, x, 6) - math.fsum((x * x).ravel())) / math.fsum((x * x).ravel()) < 1e-14
