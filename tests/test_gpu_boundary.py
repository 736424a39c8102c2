"""GPU tests of the drop-in boundary as nissa's shims call it.

* The host-buffer entry points (nq_tm_solve_eoprec_host, nq_tm_kern_square_host) take
  nissa-layout AoS host arrays: they match the oracle and give the same bits as the
  device-handle calls they wrap.
* guess_Koo of inv_tmD_cg_eoprec (cg_invert_tmD_eoprec.cpp:102-150,152-161): the start
  of the eq. 9 CG; the solution of a solve started from the previous solution's y is
  the same to the solve tolerance, reached in fewer iterations, and the guess returns
  the new y.
* Objects of a closed context refuse further use (null handle, NQ_EINVAL).
* The server replaces a batch group for a request of another kappa or precision on
  the same volume, and refuses a point source outside the lattice.
"""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import rel_l2, tm_conf, gauss_eo
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KAPPA = 0.16


@pytest.fixture(scope="module")
def lat():
    L = (8, 8, 8, 8)
    conf = tm_conf(L, 31)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    geo = orc.Geom(L)
    yield L, conf, ctx, g, geo
    ctx.close()


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_solve_host_buffers_match_device_path_and_oracle(lat, prec):
    L, conf, ctx, g, geo = lat
    src = nq.source_lx(L, 5, "gauss")
    sol_h, st_h = nq.tm_solve_eoprec_host(ctx, g, KAPPA, 0.01, 10000, 1e-24, src, prec=prec)
    assert st_h["rc"] == nq.NQ_OK and st_h["check_rel_res_sq"] < 1e-23, st_h
    fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
    fx = ctx.field(nq.SPINCOLOR, nq.LX)
    st_d = nq.tm_solve_eoprec(ctx, g, KAPPA, 0.01, 10000, 1e-24, fx, fs, prec=prec)
    assert st_d["iters"] == st_h["iters"]
    assert np.array_equal(fx.download(), sol_h)
    ref, _ = geo.tm_solve_eoprec(conf, KAPPA, 0.01, 20000, 1e-24, src)
    assert rel_l2(sol_h, ref) < 1e-10


def test_kern_square_host_buffers(lat):
    L, conf, ctx, g, geo = lat
    ce, co = geo.split(conf)
    x = gauss_eo(geo.Vh, 8)
    for mu in (0.01, -0.3):
        out = nq.tm_kern_square_host(ctx, g, KAPPA, mu, x)
        assert rel_l2(out, geo.tmDkern_square(ce, co, KAPPA, mu, x)) < 1e-14
        fi = ctx.field(nq.SPINCOLOR, nq.ODD, x)
        fo = ctx.field(nq.SPINCOLOR, nq.ODD)
        nq.tm_kern_square(ctx, g, KAPPA, mu, fo, fi)
        assert np.array_equal(fo.download(), out)
    with pytest.raises(nq.NqError):
        nq.tm_kern_square_host(ctx, g, KAPPA, 0.01, x[:-1])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_guess_Koo_device_and_host(lat, prec):
    """A first solve at mu returns y in the guess; a solve at a nearby mu started from
    it needs fewer iterations and gives the oracle's solution; a guess that already
    solves the system returns at once with the same solution."""
    L, conf, ctx, g, geo = lat
    src = nq.source_lx(L, 6, "gauss")
    fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
    fx = ctx.field(nq.SPINCOLOR, nq.LX)
    guess = ctx.field(nq.SPINCOLOR, nq.ODD, np.zeros((geo.Vh, 4, 3, 2)))
    st0 = nq.tm_solve_eoprec(ctx, g, KAPPA, 0.02, 10000, 1e-24, fx, fs, prec=prec, guess=guess)
    assert st0["rc"] == nq.NQ_OK and st0["check_rel_res_sq"] < 1e-23
    y0 = guess.download()
    assert np.linalg.norm(y0) > 0
    # the returned y solves K K^+ y = phi: K(-mu) y is the odd part of the solution (eq. 10)
    ce, co = geo.split(conf)
    xe, xo = geo.split(fx.download())
    assert rel_l2(geo.tmDkern(ce, co, KAPPA, -0.02, y0), xo) < 1e-12
    st1 = nq.tm_solve_eoprec(ctx, g, KAPPA, 0.021, 10000, 1e-24, fx, fs, prec=prec, guess=guess)
    assert st1["rc"] == nq.NQ_OK and st1["check_rel_res_sq"] < 1e-23
    assert st1["iters"] < st0["iters"], (st0["iters"], st1["iters"])
    ref, _ = geo.tm_solve_eoprec(conf, KAPPA, 0.021, 20000, 1e-24, src)
    assert rel_l2(fx.download(), ref) < 1e-10
    # a guess that is already the solution: no iteration, the same solution
    y1 = guess.download()
    st2 = nq.tm_solve_eoprec(ctx, g, KAPPA, 0.021, 10000, 1e-20, fx, fs, prec=prec, guess=guess)
    assert st2["rc"] == nq.NQ_OK and st2["iters"] == 0, st2
    assert rel_l2(fx.download(), ref) < 1e-10
    # the host entry point updates its guess buffer in place the same way
    gh = np.ascontiguousarray(y0)
    sol_h, st_h = nq.tm_solve_eoprec_host(ctx, g, KAPPA, 0.021, 10000, 1e-24, src, prec=prec, guess=gh)
    assert st_h["iters"] == st1["iters"]
    assert rel_l2(sol_h, ref) < 1e-10
    # y itself is fixed only to the CG's tolerance times the condition of K K^+: compare
    # K(-mu) y, the odd part of the solution (eq. 10)
    _, ref_o = geo.split(ref)
    assert rel_l2(geo.tmDkern(ce, co, KAPPA, -0.021, gh), ref_o) < 1e-10
    assert rel_l2(geo.tmDkern(ce, co, KAPPA, -0.021, y1), ref_o) < 1e-10
    # batched solves take a guess per source (None: start from zero)
    sols = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in range(2)]
    gb = ctx.field(nq.SPINCOLOR, nq.ODD, y0)
    res = nq.tm_solve_eoprec_multi(ctx, g, KAPPA, [0.021, 0.021], 10000, 1e-24, sols, [fs, fs], prec=prec,
                                   guesses=[gb, None])
    assert all(r["rc"] == nq.NQ_OK and r["converged"] for r in res), res
    assert res[0]["iters"] < res[1]["iters"]
    for s in sols:
        assert rel_l2(s.download(), ref) < 1e-10
    assert rel_l2(geo.tmDkern(ce, co, KAPPA, -0.021, gb.download()), ref_o) < 1e-10


def test_closed_context_refuses_its_objects():
    L = (4, 4, 4, 4)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=tm_conf(L, 3))
    f = ctx.field(nq.SPINCOLOR, nq.ODD, gauss_eo(128, 1))
    o = ctx.field(nq.SPINCOLOR, nq.ODD)
    ctx.close()
    for call in (f.download, lambda: nq.tm_kern(ctx, g, KAPPA, 0.1, o, f), g.device_bytes):
        try:
            r = call()
        except nq.NqError as e:
            assert e.code == nq.NQ_EINVAL
        else:
            assert r == 0  # device_bytes of a null handle
    f.close()
    o.close()
    g.close()


def test_server_replaces_groups_for_other_kappa_and_precision():
    """max_groups = 1: requests with two kappas and two precisions on one volume all
    finish (the one group is drained and replaced in turn) and match single solves."""
    L = (4, 6, 4, 8)
    conf = tm_conf(L, 17)
    srv = nq.Server(max_groups=1)
    vol = srv.add_volume(L, conf)
    geo = orc.Geom(L)
    plan = [(0.16, "f32", 0.01), (0.15, "f32", 0.02), (0.16, "f16", 0.005), (0.15, "f64", 0.01),
            (0.16, "f32", 0.003)]
    srcs = [nq.source_lx(L, 40 + i, "gauss") for i in range(len(plan))]
    sols = [np.zeros((geo.V, 4, 3, 2)) for _ in plan]
    ids = [srv.submit(vol, k, mu, 1e-22, prec=p, srcs=[s], sols=[x]) for (k, p, mu), s, x in zip(plan, srcs, sols)]
    for i in ids:
        r = srv.wait(i, timeout=120.0)
        assert r["status"] == nq.NQ_OK and r["max_rel_res_sq"] < 1e-22, r
    for (k, p, mu), s, x in zip(plan, srcs, sols):
        ref, _ = geo.tm_solve_eoprec(conf, k, mu, 20000, 1e-24, s)
        assert rel_l2(x, ref) < 1e-9
    assert srv.device_bytes() > 0
    srv.close()


def test_server_point_source_bounds():
    L = (4, 6, 4, 8)
    V = int(np.prod(L))
    srv = nq.Server(max_groups=1)
    vol = srv.add_volume(L, tm_conf(L, 18))
    for bad in (-1, V, V + 5):
        with pytest.raises(nq.NqError) as e:
            srv.submit(vol, KAPPA, 0.01, 1e-20, prec="f32", point_site=bad, nsrc=1)
        assert e.value.code == nq.NQ_EINVAL
    # a point source away from the origin against the host generator's point source
    site, sc = V - 3, 7
    sol = [np.zeros((V, 4, 3, 2))]
    r = srv.wait(srv.submit(vol, KAPPA, 0.01, 1e-24, prec="f64", point_site=site, nsrc=1, sols=sol))
    assert r["status"] == nq.NQ_OK
    sol7 = [np.zeros((V, 4, 3, 2)) for _ in range(8)]
    r = srv.wait(srv.submit(vol, KAPPA, 0.01, 1e-24, prec="f64", point_site=site, nsrc=8, sols=sol7))
    assert r["status"] == nq.NQ_OK
    geo = orc.Geom(L)
    src = nq.source_lx(L, 0, "point", site0=site, sc=sc)
    ref, _ = geo.tm_solve_eoprec(tm_conf(L, 18), KAPPA, 0.01, 20000, 1e-24, src)
    assert rel_l2(sol7[sc], ref) < 1e-10
    src0 = nq.source_lx(L, 0, "point", site0=site, sc=0)
    ref0, _ = geo.tm_solve_eoprec(tm_conf(L, 18), KAPPA, 0.01, 20000, 1e-24, src0)
    assert rel_l2(sol[0], ref0) < 1e-10
    srv.close()


def test_gauge_handle_matches_its_host_conf():
    """nq_gauge_matches_lx: true for the conf the handle was made from (also a copy at
    another address), false once a boundary phase has been applied in place; the
    handle rebuilt from the changed conf gives the changed operator."""
    L = (4, 6, 4, 8)
    conf = nq.gauge_random_lx(L, 19, 0.3)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    assert nq.gauge_matches_lx(g, conf) and nq.gauge_matches_lx(g, conf.copy())
    nq.gauge_phase_lx(L, conf, [1.0, 0, 0, 0])
    assert not nq.gauge_matches_lx(g, conf)
    g2 = nq.Gauge(ctx, lx=conf)
    assert nq.gauge_matches_lx(g2, conf)
    geo = orc.Geom(L)
    ce, co = geo.split(conf)
    x = gauss_eo(geo.Vh, 4)
    fi = ctx.field(nq.SPINCOLOR, nq.EVEN, x)
    fo = ctx.field(nq.SPINCOLOR, nq.ODD)
    nq.tm_hop(ctx, g2, fo, fi)
    assert rel_l2(fo.download(), geo.tm_hop(1, ce, co, x)) < 1e-14
    ctx.close()
