"""GPU parity of the staggered operators, staggered CG / multi-shift CGM, the
legacy lx CG on Q^2 and the multi-source eo solve, against the oracle through
the C ABI.

Bars: operator applies match the FP64 oracle to <= 1e-14 relative L2; solves
reach the requested FP64 true residual and agree with the oracle's solution to
1e-10 relative L2 when both run to residue 1e-24; a solve gives the same bits
alone and inside a batch of other sources (per-source arithmetic and
reductions do not depend on the batch).
"""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import gauss_eo, rel_l2, st_conf, tm_conf
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ST_SHAPES = [(8, 8, 8, 8), (4, 6, 4, 8), (6, 4, 8, 2)]


@pytest.fixture(scope="module")
def st_setups():
    cache = {}

    def get(L, seed=1):
        key = (L, seed)
        if key not in cache:
            conf = st_conf(L, seed)
            ctx = nq.Ctx(L)
            g = nq.Gauge(ctx, lx=conf)
            geo = orc.Geom(L)
            ce, co = geo.split(conf)
            cache[key] = (ctx, g, geo, conf, ce, co)
        return cache[key]

    return get


@pytest.mark.parametrize("L", ST_SHAPES)
def test_st_operators(st_setups, L):
    ctx, g, geo, conf, ce, co = st_setups(L)
    xe = gauss_eo(geo.Vh, 11, ncomp=3)
    xo = gauss_eo(geo.Vh, 12, ncomp=3)
    fe = ctx.field(nq.COLOR, nq.EVEN, xe)
    fo = ctx.field(nq.COLOR, nq.ODD, xo)
    oo = ctx.field(nq.COLOR, nq.ODD)
    oe = ctx.field(nq.COLOR, nq.EVEN)
    nq.st_2Doe(ctx, g, oo, fe)
    assert rel_l2(oo.download(), geo.st_2Doe(ce, co, xe)) < 1e-14
    nq.st_Deo_half(ctx, g, oe, fo)
    assert rel_l2(oe.download(), geo.st_Deo_half(ce, co, xo)) < 1e-14
    for m2 in (0.0, 0.01, 0.37):
        nq.st_DDee(ctx, g, m2, oe, fe)
        assert rel_l2(oe.download(), geo.stD2ee_m2(ce, co, m2, xe)) < 1e-14


def test_st_DDee_is_hermitian_positive(st_setups):
    """<y, A x> = <A y, x> and <x, A x> >= m2 |x|^2 on the device operator."""
    L = (8, 8, 8, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    x = gauss_eo(geo.Vh, 21, ncomp=3)
    y = gauss_eo(geo.Vh, 22, ncomp=3)
    fx, fy = ctx.field(nq.COLOR, nq.EVEN, x), ctx.field(nq.COLOR, nq.EVEN, y)
    ax, ay = ctx.field(nq.COLOR, nq.EVEN), ctx.field(nq.COLOR, nq.EVEN)
    m2 = 0.04
    nq.st_DDee(ctx, g, m2, ax, fx)
    nq.st_DDee(ctx, g, m2, ay, fy)
    cx = lambda a: a[..., 0] + 1j * a[..., 1]
    lhs = np.vdot(cx(y), cx(ax.download()))
    rhs = np.vdot(cx(ay.download()), cx(x))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    assert np.vdot(cx(x), cx(ax.download())).real >= m2 * np.vdot(cx(x), cx(x)).real


@pytest.mark.parametrize("L", ST_SHAPES[:2])
def test_st_cg_matches_oracle(st_setups, L):
    ctx, g, geo, conf, ce, co = st_setups(L)
    src = gauss_eo(geo.Vh, 31, ncomp=3)
    m2 = 0.0025
    ref, rst = geo.st_cg(ce, co, m2, 10000, 1e-24, src)
    fb = ctx.field(nq.COLOR, nq.EVEN, src)
    fx = ctx.field(nq.COLOR, nq.EVEN)
    st = nq.st_cg(ctx, g, m2, 10000, 1e-24, fx, fb)
    assert st["converged"] and st["rel_res_sq"] < 2e-24
    assert rel_l2(fx.download(), ref) < 1e-10
    assert abs(st["iters"] - rst["iters"]) <= 2
    # warm start from the solution converges at once
    st2 = nq.st_cg(ctx, g, m2, 10000, 1e-20, fx, fb, guess=fx)
    assert st2["converged"] and st2["iters"] <= 2


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_st_cgm_matches_oracle(st_setups, prec):
    """Multi-shift CG (cgm_invert with ApplyStD2eeM2) on 8^4, three sources and
    four shifts batched in one call; every (source, shift) reaches its residue in
    FP64 and agrees with the oracle's CGM solution."""
    L = (8, 8, 8, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    shifts = [0.0004, 0.003, 0.02, 0.25]
    residues = [1e-24] * 4
    srcs = [gauss_eo(geo.Vh, 40 + i, ncomp=3) for i in range(3)]
    fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in srcs]
    fsol = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
    res = nq.st_cgm(ctx, g, shifts, 20000, residues, fsol, fsrc, prec=prec)
    for i, s in enumerate(srcs):
        assert res[i]["rc"] == nq.NQ_OK
        ref, info = geo.st_cgm(ce, co, shifts, 20000, residues, s)
        for k in range(len(shifts)):
            assert res[i]["true_res_sq"][k] < 1e-23
            assert rel_l2(fsol[i][k].download(), ref[k]) < 1e-10, (i, k)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_st_cgm_batch_equals_alone(st_setups, prec):
    """A source solved alone and inside a batch of three gives identical bits (in
    mixed precision also through the refinement of the systems that need it)."""
    L = (4, 6, 4, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    shifts = [0.001, 0.05]
    srcs = [gauss_eo(geo.Vh, 60 + i, ncomp=3) for i in range(3)]
    fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in srcs]
    batch = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
    nq.st_cgm(ctx, g, shifts, 20000, [1e-22, 1e-22], batch, fsrc, prec=prec)
    alone = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts]]
    nq.st_cgm(ctx, g, shifts, 20000, [1e-22, 1e-22], alone, fsrc[1:2], prec=prec)
    for k in range(len(shifts)):
        assert np.array_equal(alone[0][k].download(), batch[1][k].download())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_st_cgm_nan_source_is_rejected(st_setups, prec):
    """A NaN in one source of a batched multi-shift solve fails the call with
    NQ_ENAN (nissa CRASHes on a NaN norm); the context then solves normally."""
    L = (4, 6, 4, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    shifts = [0.001, 0.05]
    srcs = [gauss_eo(geo.Vh, 70 + i, ncomp=3) for i in range(2)]
    srcs[1][5, 2, 1] = np.nan
    fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in srcs]
    sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
    with pytest.raises(nq.NqError) as e:
        nq.st_cgm(ctx, g, shifts, 20000, [1e-20] * 2, sols, fsrc, prec=prec)
    assert e.value.code == nq.NQ_ENAN
    res = nq.st_cgm(ctx, g, shifts, 20000, [1e-20] * 2, sols[:1], fsrc[:1], prec=prec)
    assert res[0]["rc"] == nq.NQ_OK and max(res[0]["true_res_sq"]) < 2e-20


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_st_zero_source_is_rejected(st_setups, prec):
    """A source of norm zero in a multi-shift solve, and in the staggered CG, is
    refused with NQ_ENAN (nissa CRASHes with "invalid norm")."""
    L = (4, 6, 4, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    shifts = [0.001, 0.05]
    srcs = [gauss_eo(geo.Vh, 80, ncomp=3), np.zeros((geo.Vh, 3, 2))]
    fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in srcs]
    sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
    with pytest.raises(nq.NqError) as e:
        nq.st_cgm(ctx, g, shifts, 20000, [1e-20] * 2, sols, fsrc, prec=prec)
    assert e.value.code == nq.NQ_ENAN
    if prec == "f64":
        x = ctx.field(nq.COLOR, nq.EVEN)
        with pytest.raises(nq.NqError) as e:
            nq.st_cg(ctx, g, 0.01, 1000, 1e-20, x, fsrc[1])
        assert e.value.code == nq.NQ_ENAN


def test_tmQ2_cg_matches_oracle():
    L = (4, 6, 4, 8)
    conf = tm_conf(L, 3)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    geo = orc.Geom(L)
    src = nq.source_lx(L, 5, "gauss")
    kappa, mu = 0.16, 0.02
    ref, rst = geo.tmQ2_cg(conf, kappa, mu, 10000, 1e-24, src)
    fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
    fx = ctx.field(nq.SPINCOLOR, nq.LX)
    st = nq.tmQ2_cg(ctx, g, kappa, mu, 10000, 1e-24, fx, fs)
    assert st["converged"] and st["rel_res_sq"] < 2e-24
    assert rel_l2(fx.download(), ref) < 1e-10
    assert abs(st["iters"] - rst["iters"]) <= 2
    ctx.close()


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_multi_rhs_eo_solve(prec):
    """Four sources with four twisted masses solved in one batched call: each
    solution matches the oracle's eo solve, and the one solved alone through the
    same entry point gives the same bits as inside the batch."""
    L = (8, 8, 8, 8)
    conf = tm_conf(L, 4)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    geo = orc.Geom(L)
    kappa = 0.16
    mus = [0.003, 0.005, 0.01, 0.02]
    srcs = [nq.source_lx(L, 70 + i, "z2", parity_mask=1) for i in range(4)]
    fsrc = [ctx.field(nq.SPINCOLOR, nq.LX, s) for s in srcs]
    fsol = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in srcs]
    res = nq.tm_solve_eoprec_multi(ctx, g, kappa, mus, 20000, 1e-24, fsol, fsrc, prec=prec)
    for i in range(4):
        assert res[i]["rc"] == nq.NQ_OK and res[i]["converged"]
        ref, _ = geo.tm_solve_eoprec(conf, kappa, mus[i], 20000, 1e-24, srcs[i])
        assert rel_l2(fsol[i].download(), ref) < 1e-10, i
    alone = [ctx.field(nq.SPINCOLOR, nq.LX)]
    nq.tm_solve_eoprec_multi(ctx, g, kappa, mus[2:3], 20000, 1e-24, alone, fsrc[2:3], prec=prec)
    assert np.array_equal(alone[0].download(), fsol[2].download())
    ctx.close()


def test_multi_rhs_graph_replays_match_eager(monkeypatch):
    """The batched solve replays captured iteration chunks (HIP graphs) between its
    eager reliable updates and polls; every source must match the eager launch
    sequence bit for bit."""
    L = (8, 8, 8, 8)
    conf = tm_conf(L, 9)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    mus = [0.003, 0.01, 0.02, 0.005, 0.0075]
    srcs = [ctx.field(nq.SPINCOLOR, nq.LX, nq.source_lx(L, 90 + i, "gauss")) for i in range(5)]
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("NQ_GRAPHS", mode)
        sols = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in srcs]
        res = nq.tm_solve_eoprec_multi(ctx, g, 0.16, mus, 20000, 1e-22, sols, srcs, prec="f16")
        assert all(r["rc"] == nq.NQ_OK and r["rel_res_sq"] < 1e-22 for r in res)
        out[mode] = [s.download() for s in sols]
    for a, b in zip(out["0"], out["1"]):
        assert np.array_equal(a, b)
    ctx.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_cgm_graph_replays_match_eager(st_setups, monkeypatch, prec):
    """The multi-shift CG replays captured chunks of iterations (HIP graphs) between
    its eager polls and reliable updates; every (source, shift) solution must match
    the eager launch sequence bit for bit, and the replays run between host work."""
    L = (8, 8, 8, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    shifts = [0.001, 0.01, 0.1]
    srcs = [ctx.field(nq.COLOR, nq.EVEN, gauss_eo(geo.Vh, 140 + i, ncomp=3)) for i in range(3)]
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("NQ_GRAPHS", mode)
        sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
        res = nq.st_cgm(ctx, g, shifts, 20000, [1e-22] * 3, sols, srcs, prec=prec)
        assert all(r["rc"] == nq.NQ_OK and max(r["true_res_sq"]) < 1e-22 for r in res)
        out[mode] = [[f.download() for f in row] for row in sols]
    for ra, rb in zip(out["0"], out["1"]):
        for a, b in zip(ra, rb):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_st_rational_sum_closed_form(st_setups, prec):
    """summ_src_and_all_inv_stD2ee_m2_cgm: with the partial fractions of
    1/((x+a)(x+b)) = (1/(b-a)) (1/(x+a) - 1/(x+b)) (cons 0, poles a, b, weights
    +-1/(b-a)), chi = (A+a)^-1 (A+b)^-1 eta, so (A+b)(A+a) chi = eta up to the solve
    tolerance; and cons only adds cons*eta."""
    L = (8, 8, 8, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    a, b = 0.01, 0.2
    srcs = [gauss_eo(geo.Vh, 100 + i, ncomp=3) for i in range(2)]
    fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in srcs]
    chis = [ctx.field(nq.COLOR, nq.EVEN) for _ in srcs]
    res = nq.st_cgm_rat_sum(ctx, g, 0.0, [a, b], [1 / (b - a), -1 / (b - a)], 20000, 1e-24, chis, fsrc, prec=prec)
    assert all(r["rc"] == nq.NQ_OK for r in res)
    for s, chi in zip(srcs, chis):
        y = geo.stD2ee_m2(ce, co, a, chi.download())
        z = geo.stD2ee_m2(ce, co, b, y)
        assert rel_l2(z, s) < 1e-9
    chis2 = [ctx.field(nq.COLOR, nq.EVEN) for _ in srcs]
    nq.st_cgm_rat_sum(ctx, g, 0.75, [a, b], [1 / (b - a), -1 / (b - a)], 20000, 1e-24, chis2, fsrc, prec=prec)
    for s, c1, c2 in zip(srcs, chis, chis2):
        assert rel_l2(c2.download() - c1.download(), 0.75 * s) < 1e-9


@pytest.mark.parametrize("g5", [False, True])
def test_tmQ2_cgm_matches_oracle(g5):
    """inv_tmQ2_cgm / inv_tmDQ_cgm: multi-mass CG on Q(0)^2 + m_k^2 over lx; every mass
    reaches its residue in FP64 and matches the oracle's CGM solution."""
    L = (4, 6, 4, 8)
    conf = tm_conf(L, 21)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    geo = orc.Geom(L)
    src = nq.source_lx(L, 22, "gauss")
    masses = [0.01, 0.05, 0.2]
    res = [1e-24] * 3
    fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
    fx = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in masses]
    st = nq.tmQ2_cgm(ctx, g, 0.16, masses, 20000, res, fx, fs, g5=g5)
    assert st["rc"] == nq.NQ_OK and max(st["true_res_sq"]) < 2e-24
    osrc = src.copy()
    if g5:
        osrc[:, 2:] *= -1
    ref, info = geo.tmQ2_cgm(conf, 0.16, masses, 20000, res, osrc)
    for k in range(3):
        assert rel_l2(fx[k].download(), ref[k]) < 1e-10, k
    ctx.close()


@pytest.mark.parametrize("L", ST_SHAPES[:2])
def test_st_apply_D_matches_oracle(st_setups, L):
    """evn/odd_apply_stD, their _dag and apply_stD (dirac_operator_stD.hpp:16-41,
    254-291) against the oracle; D^+(m) = -D(-m) gives the dagger reference."""
    ctx, g, geo, conf, ce, co = st_setups(L)
    xe = gauss_eo(geo.Vh, 71, ncomp=3)
    xo = gauss_eo(geo.Vh, 72, ncomp=3)
    fe, fo = ctx.field(nq.COLOR, nq.EVEN, xe), ctx.field(nq.COLOR, nq.ODD, xo)
    oe, oo = ctx.field(nq.COLOR, nq.EVEN), ctx.field(nq.COLOR, nq.ODD)
    m = 0.07
    re, ro = geo.apply_stD(ce, co, m, xe, xo)
    nq.st_apply_D(ctx, g, m, oe, oo, fe, fo)
    assert rel_l2(oe.download(), re) < 1e-14 and rel_l2(oo.download(), ro) < 1e-14
    de, do = geo.apply_stD(ce, co, -m, xe, xo)
    nq.st_apply_D(ctx, g, m, oe, oo, fe, fo, dag=True)
    assert rel_l2(oe.download(), -de) < 1e-14 and rel_l2(oo.download(), -do) < 1e-14
    # one parity only: the other output is left untouched
    oo.upload(np.zeros((geo.Vh, 3, 2)))
    nq.st_apply_D(ctx, g, m, oe, None, fe, fo)
    assert rel_l2(oe.download(), re) < 1e-14 and not oo.download().any()
    nq.st_apply_D(ctx, g, m, None, oo, fe, fo, dag=True)
    assert rel_l2(oo.download(), -do) < 1e-14


@pytest.mark.parametrize("L", ST_SHAPES[:2])
def test_st_solve_D_matches_oracle(st_setups, L):
    """inv_evn_stD_cg and inv_stD_cg (cg_invert_evn_stD.hpp:17-31,
    cg_invert_stD.hpp:14-88): the even solve of (m^2 - 1/4 Deo Doe) on D^+ src, the
    rebuilt odd part, and the full residual |D sol - src|^2/|src|^2."""
    ctx, g, geo, conf, ce, co = st_setups(L)
    m = 0.05
    se = gauss_eo(geo.Vh, 81, ncomp=3)
    so = gauss_eo(geo.Vh, 82, ncomp=3)
    de, _ = geo.apply_stD(ce, co, -m, se, so)
    ref_e, _ = geo.st_cg(ce, co, m * m, 10000, 1e-24, -de)
    ref_o = (so - 0.5 * geo.st_2Doe(ce, co, ref_e)) / m
    fe, fo = ctx.field(nq.COLOR, nq.EVEN, se), ctx.field(nq.COLOR, nq.ODD, so)
    xe, xo = ctx.field(nq.COLOR, nq.EVEN), ctx.field(nq.COLOR, nq.ODD)
    st = nq.st_solve_D(ctx, g, m, 10000, 1e-24, xe, xo, fe, fo)
    assert st["converged"] and st["rel_res_sq"] < 2e-24
    assert rel_l2(xe.download(), ref_e) < 1e-10 and rel_l2(xo.download(), ref_o) < 1e-10
    assert st["check_rel_res_sq"] < 1e-21
    # the solution solves D x = src in the oracle's own operator
    ae, ao = geo.apply_stD(ce, co, m, xe.download(), xo.download())
    r2 = np.sum((ae - se) ** 2) + np.sum((ao - so) ** 2)
    assert abs(r2 / (np.sum(se ** 2) + np.sum(so ** 2)) - st["check_rel_res_sq"]) < 1e-3 * st["check_rel_res_sq"] + 1e-30
    # even-only solve (inv_evn_stD_cg) gives the same even bits
    ye = ctx.field(nq.COLOR, nq.EVEN)
    st2 = nq.st_solve_D(ctx, g, m, 10000, 1e-24, ye, None, fe, fo)
    assert st2["converged"] and np.array_equal(ye.download(), xe.download())


def test_limits_are_refused_and_wide_batches_stream(st_setups):
    """The documented limits: lattice extents must be even and >= 2 (NQ_EINVAL, as
    nissa's geometry CRASHes on an odd extent), at most 16 shifts per multi-shift
    solve and 16 masses per multi-mass CG (NQ_EINVAL); and more sources than one
    launch takes (17 here, against 16) stream through in groups, each source with
    the bits it has alone."""
    for bad in [(3, 4, 4, 4), (4, 4, 4, 1), (4, 0, 4, 4)]:
        with pytest.raises(nq.NqError) as e:
            nq.Ctx(bad)
        assert e.value.code == nq.NQ_EINVAL, bad
    L = (4, 6, 4, 8)
    ctx, g, geo, conf, ce, co = st_setups(L)
    src = ctx.field(nq.COLOR, nq.EVEN, gauss_eo(geo.Vh, 90, ncomp=3))
    shifts = [0.01 * (k + 1) for k in range(17)]
    sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts]]
    with pytest.raises(nq.NqError) as e:
        nq.st_cgm(ctx, g, shifts, 1000, [1e-20] * 17, sols, [src])
    assert e.value.code == nq.NQ_EINVAL
    tconf = tm_conf(L, 91)
    gt = nq.Gauge(ctx, lx=tconf)
    fs = ctx.field(nq.SPINCOLOR, nq.LX, nq.source_lx(L, 92, "gauss"))
    masses = [0.01 * (k + 1) for k in range(17)]
    xs = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in masses]
    with pytest.raises(nq.NqError) as e:
        nq.tmQ2_cgm(ctx, gt, 0.16, masses, 1000, [1e-20] * 17, xs, fs)
    assert e.value.code == nq.NQ_EINVAL
    srcs = [ctx.field(nq.COLOR, nq.EVEN, gauss_eo(geo.Vh, 100 + i, ncomp=3)) for i in range(17)]
    sh = [0.001, 0.05]
    wide = [[ctx.field(nq.COLOR, nq.EVEN) for _ in sh] for _ in srcs]
    res = nq.st_cgm(ctx, g, sh, 20000, [1e-22] * 2, wide, srcs)
    assert all(r["rc"] == nq.NQ_OK and max(r["true_res_sq"]) < 2e-22 for r in res)
    for i in (0, 16):
        alone = [[ctx.field(nq.COLOR, nq.EVEN) for _ in sh]]
        nq.st_cgm(ctx, g, sh, 20000, [1e-22] * 2, alone, srcs[i:i + 1])
        for k in range(2):
            assert np.array_equal(alone[0][k].download(), wide[i][k].download()), (i, k)
