"""The multi-shift CG with each chunk of iterations in one cooperative launch
(NQ_CGM_PERSIST=1, launch_cgm_persistent): the five phases of an iteration run in
one kernel with grid barriers between them, on the same block code and logical
blocks as the separate launches, so every (source, shift) solution must match the
eager launch sequence bit for bit; and a barrier that waits past its limit ends the
launch with an error instead of a wait (the limit forced from the host).
"""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import gauss_eo, st_conf
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

L = (8, 8, 8, 8)


@pytest.fixture(scope="module")
def setup():
    conf = st_conf(L, 1)
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    geo = orc.Geom(L)
    yield ctx, g, geo
    g.close()
    ctx.close()


def solve(ctx, g, geo, nsrc, prec, shifts, seed=150):
    srcs = [ctx.field(nq.COLOR, nq.EVEN, gauss_eo(geo.Vh, seed + i, ncomp=3)) for i in range(nsrc)]
    sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
    res = nq.st_cgm(ctx, g, shifts, 20000, [1e-22] * len(shifts), sols, srcs, prec=prec)
    return res, [[f.download() for f in row] for row in sols]


# 3 sources: one group of register-blocked sources; 8 FP32 sources: RB = 8
@pytest.mark.parametrize("prec,nsrc", [("f64", 3), ("f32", 3), ("f32", 8), ("f64", 5)])
def test_cgm_persistent_matches_eager(setup, monkeypatch, prec, nsrc):
    ctx, g, geo = setup
    shifts = [0.001, 0.01, 0.1]
    monkeypatch.setenv("NQ_GRAPHS", "0")
    monkeypatch.delenv("NQ_CGM_PERSIST", raising=False)
    res_e, eager = solve(ctx, g, geo, nsrc, prec, shifts)
    monkeypatch.setenv("NQ_CGM_PERSIST", "1")
    res_p, pers = solve(ctx, g, geo, nsrc, prec, shifts)
    for re_, rp in zip(res_e, res_p):
        assert re_["rc"] == nq.NQ_OK and rp["rc"] == nq.NQ_OK
        assert max(rp["true_res_sq"]) < 1e-22
        assert re_["iters"] == rp["iters"]
    for ra, rb in zip(eager, pers):
        for a, b in zip(ra, rb):
            assert np.array_equal(a, b)


def test_cgm_persistent_barrier_limit(setup, monkeypatch):
    """A barrier wait limit of zero polls: the first workgroup that has to wait
    raises the abort word, every workgroup leaves the launch, and the solve
    returns NQ_EDEVICE; the context then solves normally."""
    ctx, g, geo = setup
    shifts = [0.001, 0.01, 0.1]
    monkeypatch.setenv("NQ_CGM_PERSIST", "1")
    monkeypatch.setenv("NQ_CGM_SPIN", "0")
    with pytest.raises(nq.NqError) as ei:
        solve(ctx, g, geo, 3, "f64", shifts)
    assert ei.value.code == nq.NQ_EDEVICE
    monkeypatch.delenv("NQ_CGM_SPIN")
    res, _ = solve(ctx, g, geo, 3, "f64", shifts)
    assert all(r["rc"] == nq.NQ_OK and max(r["true_res_sq"]) < 1e-22 for r in res)
