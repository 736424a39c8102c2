"""nq_tm_kern_square_noalloc (K K^+ with the library's static device scratch, no
allocation) against nq_tm_kern_square through the C ABI: bit-identical outputs on two
calls in a row, the second on a smaller lattice, both also against the oracle's
tmDkern_eoprec_square_eos; a lattice past the scratch limit is refused."""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import gauss_eo, rel_l2, tm_conf
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _apply_both(L, seed, kappa, mu, anis=None):
    conf = tm_conf(L, seed)
    geo = orc.Geom(L)
    ce, co = geo.split(conf)
    with nq.Ctx(L) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        x = gauss_eo(ctx.Vh, seed + 100)
        fi = ctx.field(nq.SPINCOLOR, nq.ODD, x)
        fa = ctx.field(nq.SPINCOLOR, nq.ODD)
        fb = ctx.field(nq.SPINCOLOR, nq.ODD)
        nq.tm_kern_square_noalloc(ctx, g, kappa, mu, fa, fi, anis=anis)
        nq.tm_kern_square(ctx, g, kappa, mu, fb, fi, anis=anis)
        a, b = fa.download(), fb.download()
    ref = geo.tmDkern_square(ce, co, kappa, mu, x, anis=anis)
    return a, b, ref


def test_noalloc_matches_allocating_entry_point():
    print("noalloc: 32x16^3 then 16x8^3", flush=True)
    a1, b1, r1 = _apply_both((32, 16, 16, 16), 21, 0.16, 0.005)
    assert np.array_equal(a1, b1)
    assert rel_l2(a1, r1) < 1e-14
    # second call: a smaller lattice, the same device array (the first call's contents stay in it)
    a2, b2, r2 = _apply_both((16, 8, 8, 8), 22, 0.15, 0.02, anis=0.9)
    assert np.array_equal(a2, b2)
    assert rel_l2(a2, r2) < 1e-14


def test_noalloc_refuses_past_the_limit():
    small = (8, 4, 4, 4)
    with nq.Ctx(small) as cs, nq.Ctx((64, 32, 32, 32)) as cb:
        g = nq.Gauge(cs, lx=tm_conf(small, 3))
        fi = cb.field(nq.SPINCOLOR, nq.ODD)
        fo = cb.field(nq.SPINCOLOR, nq.ODD)
        with pytest.raises(nq.NqError) as ei:
            nq.tm_kern_square_noalloc(cb, g, 0.16, 0.01, fo, fi)
        assert ei.value.code == nq.NQ_EINVAL and "limit" in str(ei.value)
