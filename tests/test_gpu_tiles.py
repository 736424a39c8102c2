"""Site-order tile autotuner (nq_ctx_create from 2^18 sites per parity up): the choice
lands in the tile cache, a second context of the lattice reads it, and the hop gives
the same bits in the tuned order as in the default one (the order moves sites, not the
arithmetic of a site)."""
import os

import numpy as np
import pytest

import nissa_amd as nq
from helpers import gauss_eo, tm_conf

pytestmark = pytest.mark.gpu


def test_tuned_tiles_cached_and_bit_identical(tmp_path, monkeypatch):
    L = (48, 24, 24, 24)
    cache = tmp_path / "site_tiles.txt"
    monkeypatch.setenv("NQ_TUNE_CACHE", str(cache))
    monkeypatch.delenv("NQ_TTILE", raising=False)
    monkeypatch.delenv("NQ_XYTILE", raising=False)
    monkeypatch.delenv("NQ_TUNE_TILES", raising=False)
    print("tile tuner at 48x24^3", flush=True)
    conf = tm_conf(L, 9)
    x = gauss_eo(L[0] * L[1] * L[2] * L[3] // 2, 10)
    outs = {}
    with nq.Ctx(L) as c1:
        tiles = c1.site_tiles()
        lines = [l.split() for l in cache.read_text().splitlines() if l.strip()]
        assert len(lines) == 1 and lines[0][1:5] == [str(v) for v in L], lines
        assert [int(v) for v in lines[0][5:8]] == list(tiles)
        assert L[0] % tiles[0] == 0 and L[1] % tiles[1] == 0 and L[2] % tiles[2] == 0
        g = nq.Gauge(c1, lx=conf)
        fi = c1.field(nq.SPINCOLOR, nq.ODD, x)
        fo = c1.field(nq.SPINCOLOR, nq.EVEN)
        nq.tm_hop(c1, g, fo, fi)
        outs["tuned"] = fo.download()
    with nq.Ctx(L) as c2:  # read from the cache, no second timing run
        assert c2.site_tiles() == tiles
        assert len(cache.read_text().splitlines()) == 1
    monkeypatch.setenv("NQ_TUNE_TILES", "0")
    with nq.Ctx(L) as c3:
        assert c3.site_tiles() == (48, 4, 2)
        g = nq.Gauge(c3, lx=conf)
        fi = c3.field(nq.SPINCOLOR, nq.ODD, x)
        fo = c3.field(nq.SPINCOLOR, nq.EVEN)
        nq.tm_hop(c3, g, fo, fi)
        outs["default"] = fo.download()
    assert np.array_equal(outs["tuned"], outs["default"])
