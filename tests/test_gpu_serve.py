"""GPU tests of the propagator-solve server (continuous batching over resident
gauge fields): results against the oracle, the same bits alone and in a mix of
other requests, admission by HBM footprint, and refusal when nothing fits.
"""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import rel_l2, tm_conf
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KAPPA = 0.16


def point_sols(L, n):
    V = int(np.prod(L))
    return [np.zeros((V, 4, 3, 2)) for _ in range(n)]


def test_server_matches_oracle():
    L = (4, 6, 4, 8)
    conf = tm_conf(L, 11)
    srv = nq.Server(max_groups=1)
    vol = srv.add_volume(L, conf)
    srcs = [nq.source_lx(L, 80 + i, "gauss") for i in range(3)]
    sols = point_sols(L, 3)
    rid = srv.submit(vol, KAPPA, 0.01, 1e-24, prec="f32", srcs=srcs, sols=sols)
    res = srv.wait(rid)
    assert res["status"] == nq.NQ_OK and res["max_rel_res_sq"] < 1e-23
    geo = orc.Geom(L)
    for s, x in zip(srcs, sols):
        ref, _ = geo.tm_solve_eoprec(conf, KAPPA, 0.01, 20000, 1e-24, s)
        assert rel_l2(x, ref) < 1e-10
    srv.close()


def test_server_point_sources_and_mix_equality():
    """A request solved alone and again inside a burst of requests on two volumes
    with other masses gives bit-identical solutions."""
    La, Lb = (8, 8, 8, 8), (4, 6, 4, 8)
    ca, cb = tm_conf(La, 12), tm_conf(Lb, 13)
    srv = nq.Server(max_groups=2)
    va = srv.add_volume(La, ca)
    vb = srv.add_volume(Lb, cb)
    alone = point_sols(La, 12)
    rid = srv.submit(va, KAPPA, 0.005, 1e-20, prec="f16", point_site=0, sols=alone)
    r0 = srv.wait(rid)
    assert r0["status"] == nq.NQ_OK and r0["max_rel_res_sq"] < 1e-20
    # the point source solution of spin-colour 5 against the oracle
    geo = orc.Geom(La)
    src5 = nq.source_lx(La, 0, "point", site0=0, sc=5)
    ref, _ = geo.tm_solve_eoprec(ca, KAPPA, 0.005, 20000, 1e-24, src5)
    assert rel_l2(alone[5], ref) < 1e-9
    # burst: other masses, another volume, and the same request again
    mixed = point_sols(La, 12)
    ids = [
        srv.submit(va, KAPPA, 0.02, 1e-20, prec="f16", point_site=17),
        srv.submit(vb, KAPPA, 0.003, 1e-20, prec="f16", point_site=3),
        srv.submit(va, KAPPA, 0.01, 1e-20, prec="f16", point_site=5, nsrc=7),
        srv.submit(va, KAPPA, 0.005, 1e-20, prec="f16", point_site=0, sols=mixed),
        srv.submit(vb, KAPPA, 0.02, 1e-20, prec="f16", point_site=9),
    ]
    for i in ids:
        r = srv.wait(i)
        assert r["status"] == nq.NQ_OK and r["max_rel_res_sq"] < 1e-20
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]
    for k in range(12):
        assert np.array_equal(alone[k], mixed[k]), k
    srv.close()


def test_server_nan_source_fails_alone():
    """A request whose source holds a NaN fails with NQ_ENAN; the requests batched
    with it in the same group still converge, with the same bits as alone."""
    L = (4, 6, 4, 8)
    conf = tm_conf(L, 16)
    good = [nq.source_lx(L, 90 + i, "gauss") for i in range(2)]
    bad = [nq.source_lx(L, 95 + i, "gauss") for i in range(2)]
    bad[1][17, 2, 1, 0] = np.nan
    srv = nq.Server(max_groups=1)
    vol = srv.add_volume(L, conf)
    alone = point_sols(L, 2)
    r = srv.wait(srv.submit(vol, KAPPA, 0.01, 1e-20, prec="f32", srcs=good, sols=alone))
    assert r["status"] == nq.NQ_OK
    mixed = point_sols(L, 2)
    ids = [srv.submit(vol, KAPPA, 0.02, 1e-20, prec="f32", srcs=bad),
           srv.submit(vol, KAPPA, 0.01, 1e-20, prec="f32", srcs=good, sols=mixed)]
    rb, rg = srv.wait(ids[0]), srv.wait(ids[1])
    assert rb["status"] == nq.NQ_ENAN, rb
    assert rg["status"] == nq.NQ_OK and rg["max_rel_res_sq"] < 1e-20, rg
    for a, m in zip(alone, mixed):
        assert np.array_equal(a, m)
    srv.close()


def test_server_admission_by_footprint():
    """With a budget that fits one batch group (64 MiB of headroom plus this tiny
    lattice's fields), 42 sources stream through its 16 slots; with a budget below
    one group the request is refused (NQ_EBUSY)."""
    L = (4, 6, 4, 8)
    conf = tm_conf(L, 14)
    srv = nq.Server(hbm_budget=100 << 20, max_groups=3)
    vol = srv.add_volume(L, conf)
    ids = [srv.submit(vol, KAPPA, mu, 1e-20, prec="f32", point_site=i) for i, mu in enumerate([0.003, 0.01, 0.02])]
    ids += [srv.submit(vol, KAPPA, 0.005, 1e-20, prec="f32", point_site=7, nsrc=6)]
    for i in ids:
        r = srv.wait(i)
        assert r["status"] == nq.NQ_OK and r["max_rel_res_sq"] < 1e-20
    assert srv.device_bytes() <= 100 << 20
    srv.close()
    tiny = nq.Server(hbm_budget=1 << 20, max_groups=1)
    v = tiny.add_volume(L, conf)
    r = tiny.wait(tiny.submit(v, KAPPA, 0.01, 1e-20, prec="f32", point_site=0, nsrc=2))
    assert r["status"] == nq.NQ_EBUSY
    tiny.close()


def test_point_source_fp16_far_field():
    """Regression: FP16 fields of a point source decay below FP32's normal range far
    from the source; such sites must store as zero (no 1/denormal overflow)."""
    L = (32, 16, 16, 16)
    conf = tm_conf(L, 15)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    fs = ctx.field(nq.SPINCOLOR, nq.LX, nq.source_lx(L, 0, "point", site0=0, sc=4))
    fx = ctx.field(nq.SPINCOLOR, nq.LX)
    st = nq.tm_solve_eoprec(ctx, g, KAPPA, 0.003, 100000, 1e-20, fx, fs, prec="f16")
    assert st["rc"] == nq.NQ_OK and st["check_rel_res_sq"] < 1e-20
    ctx.close()


def test_server_twisted_clover_requests():
    """Requests with a cSW are twisted clover solves: they match the oracle's
    inv_tmclovD_cg_eoprec, give the same bits alone and next to twisted-mass requests
    and clover requests of another mass on the same volume (separate groups)."""
    L = (8, 8, 8, 8)
    conf = tm_conf(L, 14)
    csw = 1.1
    srv = nq.Server(max_groups=3)
    vol = srv.add_volume(L, conf)
    srcs = [nq.source_lx(L, 120 + i, "gauss") for i in range(3)]
    alone = point_sols(L, 3)
    r0 = srv.wait(srv.submit(vol, KAPPA, 0.01, 1e-22, prec="f16", srcs=srcs, sols=alone, csw=csw))
    assert r0["status"] == nq.NQ_OK and r0["max_rel_res_sq"] < 1e-22, r0
    geo = orc.Geom(L)
    for s, x in zip(srcs, alone):
        ref, rst = geo.tmclov_solve_eoprec(conf, KAPPA, csw, 0.01, 20000, 1e-24, s)
        assert rst["rc"] == 0
        assert rel_l2(x, ref) < 1e-9
    mixed = point_sols(L, 3)
    ids = [
        srv.submit(vol, KAPPA, 0.02, 1e-22, prec="f16", point_site=5, csw=csw),
        srv.submit(vol, KAPPA, 0.01, 1e-22, prec="f16", point_site=9),
        srv.submit(vol, KAPPA, 0.01, 1e-22, prec="f16", srcs=srcs, sols=mixed, csw=csw),
    ]
    for i in ids:
        r = srv.wait(i)
        assert r["status"] == nq.NQ_OK and r["max_rel_res_sq"] < 1e-22, r
    srv.close()
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b)
