"""CPU-side checks of the product library: the C ABI exports what the header
declares, and the host harness helpers (gauge generator, phases, sources) do
what they promise.  No device calls are made here."""
import os
import re

import numpy as np
import pytest
import scipy.linalg

import nissa_amd as nq
from helpers import parity, cplx
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "nissa_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nq_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = nq.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert not [n for n in names if n not in nq._SIG]


def test_gauge_generator_is_su3_and_matches_expm():
    L = (4, 4, 4, 4)
    conf = nq.gauge_random_lx(L, 123, 0.3)
    U = cplx(conf)  # (V,4,3,3)
    I = np.eye(3)
    dev = np.abs(np.einsum("...ij,...kj->...ik", U, U.conj()) - I).max()
    assert dev < 1e-14
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-14
    for lx, mu in [(0, 0), (17, 2), (255, 3)]:
        H, Ul = nq.link_algebra(123, lx, mu, 0.3)
        Hc = cplx(H)
        assert np.abs(Hc - Hc.conj().T).max() < 1e-15 and abs(np.trace(Hc)) < 1e-14
        ref = scipy.linalg.expm(1j * 0.3 * Hc)
        assert np.abs(cplx(Ul) - ref).max() < 1e-14
        assert np.array_equal(Ul, conf[lx, mu])
    # determinism and seed dependence
    assert np.array_equal(conf, nq.gauge_random_lx(L, 123, 0.3))
    assert not np.array_equal(conf, nq.gauge_random_lx(L, 124, 0.3))


def test_gauge_generator_plaquette_matches_reference_construction():
    # the reference's U = exp(i 0.3 H) construction gives plaquette ~0.612 (a run of
    # nissa's own generator)
    L = (8, 8, 8, 8)
    g = orc.Geom(L)
    p = g.plaquette(nq.gauge_random_lx(L, 1, 0.3))
    assert abs(p - 0.612) < 0.003


def test_phases_and_sources():
    L = (4, 2, 2, 6)
    V = int(np.prod(L))
    conf = nq.gauge_random_lx(L, 5, 0.3)
    c2 = nq.gauge_phase_lx(L, conf.copy(), [1.0, 0, 0, 0.5])
    ph_t = np.exp(1j * np.pi / L[0])
    ph_z = np.exp(1j * np.pi * 0.5 / L[3])
    assert np.allclose(cplx(c2[:, 0]), cplx(conf[:, 0]) * ph_t, atol=1e-15)
    assert np.allclose(cplx(c2[:, 3]), cplx(conf[:, 3]) * ph_z, atol=1e-15)
    assert np.array_equal(c2[:, 1], conf[:, 1])
    s = nq.gauge_stagphases_lx(L, conf.copy())
    # eta_1 = (-1)^t : sign flips exactly where t is odd
    t = np.arange(V) // (L[1] * L[2] * L[3])
    assert np.array_equal(s[t % 2 == 1, 1], -conf[t % 2 == 1, 1])
    assert np.array_equal(s[:, 0], conf[:, 0])
    z2 = nq.source_lx(L, 3, "z2", 12, 1)
    par = parity(L)
    assert np.all(z2[par == 1] == 0)
    assert set(np.unique(z2[par == 0][..., 0])) == {-1.0, 1.0}
    assert np.all(z2[..., 1] == 0)
    pt = nq.source_lx(L, 0, "point", 12, 3, site0=5, sc=7)
    assert pt.sum() == 1.0 and pt[5].reshape(12, 2)[7, 0] == 1.0
    ga = nq.source_lx(L, 3, "gauss", 3, 3)
    assert abs(np.var(ga) - 0.5) < 0.1


def test_entry_points_fail_cleanly_without_a_device():
    """Without a usable GPU every entry point that needs one returns an error code
    (no CPU fallback, no crash): contexts, servers."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(nq.NqError) as e:
        nq.Ctx((4, 4, 4, 4))
    assert e.value.code in (nq.NQ_EDEVICE, nq.NQ_EINVAL)
    with pytest.raises(nq.NqError) as e:
        nq.Server()
    assert e.value.code == nq.NQ_EDEVICE


def test_abi_structs_match_the_header(tmp_path):
    """The ctypes mirrors of the structures crossing the ABI have the C compiler's
    layout of include/nissa_amd.h (field offsets and sizes)."""
    import ctypes as C
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    structs = {"nq_solve_stats": nq.SolveStats, "nq_serve_request": nq.ServeRequest,
               "nq_serve_result": nq.ServeResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nissa_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_closing_a_context_closes_its_objects(monkeypatch):
    """A context destroys its still-open gauge and field handles when it closes and marks
    the Python objects closed: using one afterwards passes a null handle, which the
    library refuses with NQ_EINVAL, and never the freed one.  The library is replaced by
    a recorder of the calls (no device here)."""
    import ctypes as C

    calls, live = [], set()

    class FakeLib:
        def __init__(self):
            self.next = 100

        def _new(self, out):
            self.next += 1
            out._obj.value = self.next
            live.add(self.next)
            return nq.NQ_OK

        def nq_ctx_create(self, T, X, Y, Z, dev, out):
            return self._new(out)

        def nq_field_create(self, ctx, kind, cov, out):
            return self._new(out)

        def nq_gauge_create_lx(self, ctx, conf, out):
            return self._new(out)

        def _destroy(self, name, h):
            v = h.value if isinstance(h, C.c_void_p) else h
            calls.append((name, v))
            live.discard(v)

        def nq_field_destroy(self, h):
            self._destroy("field", h)

        def nq_gauge_destroy(self, h):
            self._destroy("gauge", h)

        def nq_ctx_destroy(self, h):
            self._destroy("ctx", h)

        def nq_field_download(self, h, buf):
            v = h.value if isinstance(h, C.c_void_p) else h
            if not v:
                return nq.NQ_EINVAL
            assert v in live, "a freed handle reached the library"
            return nq.NQ_OK

        def nq_last_error(self):
            return b"null argument"

    monkeypatch.setattr(nq, "_lib", FakeLib())
    ctx = nq.Ctx((4, 4, 4, 4))
    g = nq.Gauge(ctx, lx=np.zeros((256, 4, 3, 3, 2)))
    f = ctx.field(nq.SPINCOLOR, nq.ODD)
    f.download()
    ctx.close()
    assert [c[0] for c in calls] == ["field", "gauge", "ctx"]
    assert f.h is None and g.h is None
    with pytest.raises(nq.NqError) as e:
        f.download()
    assert e.value.code == nq.NQ_EINVAL
    # closing the objects afterwards destroys nothing twice
    f.close()
    g.close()
    assert len(calls) == 3
    with pytest.raises(nq.NqError):
        nq.Gauge(ctx, lx=np.zeros((10, 4, 3, 3, 2)))


def test_conf_fingerprint_sees_in_place_changes():
    """The gauge cache's fingerprint (nq_host_conf_fingerprint): equal for equal
    content at another address, different after each of nissa's in-place changes of
    a conf (a boundary phase on any one direction, staggered phases, another conf
    read into the same buffer)."""
    L = (8, 4, 4, 6)
    conf = nq.gauge_random_lx(L, 9, 0.3)
    f0 = nq.conf_fingerprint(conf)
    assert nq.conf_fingerprint(conf.copy()) == f0
    for mu in range(4):
        th = [0.0] * 4
        th[mu] = 1.0
        assert nq.conf_fingerprint(nq.gauge_phase_lx(L, conf.copy(), th)) != f0, mu
    assert nq.conf_fingerprint(nq.gauge_stagphases_lx(L, conf.copy())) != f0
    assert nq.conf_fingerprint(nq.gauge_random_lx(L, 10, 0.3)) != f0
    # eo halves, as nq_gauge_create_eo takes them
    e, o = nq.split_lx(L, conf)
    assert nq.conf_fingerprint(e) != nq.conf_fingerprint(o)
