"""nissa_amd: MI355X-native fermion-solve path for nissa, Python bindings.

The product is the C-ABI library ``libnissa_amd.so`` (declared in
``include/nissa_amd.h``); this module only binds it with ctypes for tests,
benchmarks and scripting.  Loading fails loudly when the library has not been
built: there is no fallback implementation.

Host arrays follow nissa's CPU layout: spincolor ``(..., 4, 3, 2)`` float64,
color ``(..., 3, 2)``, quad_su3 ``(..., 4, 3, 3, 2)``; lx site order t-major,
z fastest; eo fields hold the Vh sites of one parity in lx>>1 order.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnissa_amd.so")

NQ_OK, NQ_EINVAL, NQ_ENOCONV, NQ_ENAN, NQ_EOOM, NQ_EDEVICE, NQ_EBUSY = range(7)
EVEN, ODD, LX = 0, 1, 2
SPINCOLOR, COLOR = 0, 1
PREC_F64, PREC_F32, PREC_F16, PREC_F16H = 0, 1, 2, 3
PREC = {"f64": PREC_F64, "f32": PREC_F32, "f16": PREC_F16, "f16h": PREC_F16H}


class NqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nissa_amd error {code}: {msg}")
        self.code = code


class SolveStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int),
        ("reliable_updates", C.c_int),
        ("converged", C.c_int),
        ("reserved", C.c_int),
        ("rel_res_sq", C.c_double),
        ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServeRequest(C.Structure):
    _fields_ = [
        ("volume", C.c_int),
        ("prec", C.c_int),
        ("kappa", C.c_double),
        ("mu", C.c_double),
        ("residue", C.c_double),
        ("niter", C.c_int),
        ("nsrc", C.c_int),
        ("src_kind", C.c_int),
        ("reserved", C.c_int),
        ("site0", C.c_int64),
        ("src", C.c_void_p),
        ("sol", C.c_void_p),
        ("csw", C.c_double),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("id", C.c_int64),
        ("status", C.c_int),
        ("nsrc", C.c_int),
        ("iters_max", C.c_int),
        ("reliable_updates", C.c_int),
        ("max_rel_res_sq", C.c_double),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_DP = C.POINTER(C.c_double)
_SIG = {
    "nq_clover_create": ([_P, _P, _D, C.POINTER(_P)], _I),
    "nq_clover_download_lx": ([_P, _DP], _I),
    "nq_clover_destroy": ([_P], None),
    "nq_tmclov_Dee": ([_P, _P, _D, _D, _I, _P, _P], _I),
    "nq_tmclov_inv_Dee": ([_P, _P, _D, _D, _I, _P, _P], _I),
    "nq_tmclov_kern": ([_P, _P, _P, _D, _DP, _D, _I, _P, _P], _I),
    "nq_tmclov_kern_square": ([_P, _P, _P, _D, _DP, _D, _P, _P], _I),
    "nq_tmclov_apply_Q": ([_P, _P, _P, _D, _DP, _D, _P, _P], _I),
    "nq_tmclov_apply_Q2": ([_P, _P, _P, _D, _DP, _D, _P, _P], _I),
    "nq_tmclovQ2_cg": ([_P, _P, _P, _D, _DP, _D, _I, _D, _P, _P, _P, C.POINTER(SolveStats)], _I),
    "nq_tmclovQ2_cgm": ([_P, _P, _P, _D, _DP, _I, _DP, _I, _DP, _P, _P, _P, _DP], _I),
    "nq_tmclov_cg_kernsq": ([_P, _P, _P, _D, _DP, _D, _I, _D, _I, _P, _P, _P, _P], _I),
    "nq_tmclov_solve_eoprec": ([_P, _P, _P, _D, _DP, _D, _I, _D, _I, _P, _P, _P, _DP], _I),
    "nq_tmclov_solve_eoprec_multi": ([_P, _P, _P, _D, _DP, _D, _I, _I, _D, _I, _P, _P, _P], _I),
    "nq_st_rootst_force": ([_P, _P, _DP, _DP, _I, _DP, _DP, _I, _D, _I, _P, _DP, _DP, _P], _I),
    "nq_st_apply_D": ([_P, _P, _D, _I, _P, _P, _P, _P], _I),
    "nq_st_solve_D": ([_P, _P, _D, _I, _D, _P, _P, _P, _P, _P, _P, _DP], _I),
    "nq_tmQ2_cgm": ([_P, _P, _D, _I, _DP, _I, _DP, _P, _P, _P, _DP], _I),
    "nq_tmDQ_cgm": ([_P, _P, _D, _I, _DP, _I, _DP, _P, _P, _P, _DP], _I),
    "nq_st_cgm_rat_sum": ([_P, _P, _I, _I, _D, _DP, _DP, _I, _D, _I, _P, _P, _P], _I),
    "nq_bench_tm_hop_multi": ([_P, _P, _I, _I, _I, _P, _DP, _DP], _I),
    "nq_server_create": ([_I, C.c_int64, _I, C.POINTER(_P)], _I),
    "nq_server_add_volume": ([_P, _I, _I, _I, _I, _DP, C.POINTER(_I)], _I),
    "nq_server_submit": ([_P, C.POINTER(ServeRequest), C.POINTER(C.c_int64)], _I),
    "nq_server_wait": ([_P, C.c_int64, _D, C.POINTER(ServeResult)], _I),
    "nq_server_clock": ([_P], _D),
    "nq_server_device_bytes": ([_P], C.c_int64),
    "nq_server_destroy": ([_P], None),
    "nq_last_error": ([], C.c_char_p),
    "nq_device_count": ([C.POINTER(C.c_int)], _I),
    "nq_ctx_create": ([_I, _I, _I, _I, _I, C.POINTER(_P)], _I),
    "nq_ctx_destroy": ([_P], None),
    "nq_ctx_sync": ([_P], _I),
    "nq_ctx_site_tiles": ([_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)], _I),
    "nq_ctx_set_profiling": ([_P, _I], _I),
    "nq_ctx_kernel_count": ([_P, C.POINTER(C.c_int)], _I),
    "nq_ctx_kernel_stat": ([_P, _I, C.c_char_p, _I, C.POINTER(C.c_int64), _DP, _DP], _I),
    "nq_ctx_reset_kernel_stats": ([_P], _I),
    "nq_gauge_create_eo": ([_P, _DP, _DP, C.POINTER(_P)], _I),
    "nq_gauge_create_lx": ([_P, _DP, C.POINTER(_P)], _I),
    "nq_gauge_destroy": ([_P], None),
    "nq_gauge_device_bytes": ([_P], C.c_int64),
    "nq_host_conf_fingerprint": ([_DP, C.c_int64, C.POINTER(C.c_uint64)], _I),
    "nq_gauge_matches_lx": ([_P, _DP, C.POINTER(_I)], _I),
    "nq_gauge_matches_eo": ([_P, _DP, _DP, C.POINTER(_I)], _I),
    "nq_gauge_plaquette": ([_P, _P, _DP], _I),
    "nq_read_ildg_gauge_conf": ([C.c_char_p, _I, _I, _I, _I, _DP, C.POINTER(C.c_uint32)], _I),
    "nq_host_unitarity_check_lx": ([_I, _I, _I, _I, _DP, _DP, C.POINTER(C.c_int64)], _I),
    "nq_read_lime_real_vector": ([C.c_char_p, C.c_char_p, _I, _I, _I, _I, _I, _DP, C.POINTER(C.c_uint32)], _I),
    "nq_field_create": ([_P, _I, _I, C.POINTER(_P)], _I),
    "nq_field_destroy": ([_P], None),
    "nq_field_upload": ([_P, _DP], _I),
    "nq_field_download": ([_P, _DP], _I),
    "nq_field_copy": ([_P, _P], _I),
    "nq_field_norm2": ([_P, _DP], _I),
    "nq_tm_hop": ([_P, _P, _DP, _P, _P], _I),
    "nq_tm_Dee": ([_P, _D, _D, _I, _P, _P], _I),
    "nq_tm_kern": ([_P, _P, _D, _DP, _D, _P, _P], _I),
    "nq_tm_kern_square": ([_P, _P, _D, _DP, _D, _P, _P], _I),
    "nq_tm_kern_square_noalloc": ([_P, _P, _D, _DP, _D, _P, _P], _I),
    "nq_tm_apply_Q": ([_P, _P, _D, _DP, _D, _P, _P], _I),
    "nq_tm_apply_Q2": ([_P, _P, _D, _D, _P, _P], _I),
    "nq_tm_cg_kernsq": ([_P, _P, _D, _DP, _D, _I, _D, _I, _P, _P, _P, C.POINTER(SolveStats)], _I),
    "nq_tm_solve_eoprec": ([_P, _P, _D, _DP, _D, _I, _D, _I, _P, _P, _P, C.POINTER(SolveStats), _DP], _I),
    "nq_tmQ2_cg": ([_P, _P, _D, _D, _I, _D, _P, _P, _P, C.POINTER(SolveStats)], _I),
    "nq_tm_solve_eoprec_multi": ([_P, _P, _D, _DP, _I, _DP, _I, _D, _I, C.POINTER(_P), C.POINTER(_P),
                                  C.POINTER(_P), C.POINTER(SolveStats)], _I),
    "nq_st_2Doe": ([_P, _P, _P, _P], _I),
    "nq_st_Deo_half": ([_P, _P, _P, _P], _I),
    "nq_st_DDee": ([_P, _P, _D, _P, _P], _I),
    "nq_st_cg": ([_P, _P, _D, _I, _D, _P, _P, _P, C.POINTER(SolveStats)], _I),
    "nq_st_cgm": ([_P, _P, _I, _I, _DP, _I, _DP, _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(SolveStats), _DP], _I),
    "nq_tm_kern_square_host": ([_P, _P, _D, _DP, _D, _DP, _DP], _I),
    "nq_tm_solve_eoprec_host": ([_P, _P, _D, _DP, _D, _I, _D, _I, _DP, _DP, _DP, C.POINTER(SolveStats), _DP],
                                _I),
    "nq_host_gauge_random_lx": ([_I, _I, _I, _I, C.c_uint64, _D, _DP], _I),
    "nq_host_link_algebra": ([C.c_uint64, C.c_int64, _I, _D, _DP, _DP], _I),
    "nq_host_gauge_phase_lx": ([_I, _I, _I, _I, _DP, _DP], _I),
    "nq_host_gauge_stagphases_lx": ([_I, _I, _I, _I, _DP], _I),
    "nq_host_source_lx": ([_I, _I, _I, _I, C.c_uint64, _I, _I, _I, C.c_int64, _I, _DP], _I),
    "nq_host_split_lx": ([_I, _I, _I, _I, _I, _DP, _DP, _DP], _I),
    "nq_host_paste_lx": ([_I, _I, _I, _I, _I, _DP, _DP, _DP], _I),
    "nq_bench_tm_hop": ([_P, _P, _I, _I, _P, _P, _DP, _DP], _I),
}


def lib():
    """Load libnissa_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build() (no fallback path exists)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIG.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return {n: hasattr(L, n) for n in _SIG}


def _check(rc):
    if rc != NQ_OK:
        raise NqError(rc, lib().nq_last_error().decode())
    return rc


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def _anis(anis):
    return None if anis is None else C.byref(C.c_double(anis))


def device_count():
    n = C.c_int(0)
    rc = lib().nq_device_count(C.byref(n))
    return n.value if rc == NQ_OK else 0


# ----------------------------------------------------------------- host helpers
def dims(L):
    T, X, Y, Z = L
    return int(T), int(X), int(Y), int(Z)


def volume(L):
    T, X, Y, Z = dims(L)
    return T * X * Y * Z


def gauge_random_lx(L, seed, eps=0.3):
    V = volume(L)
    conf = np.empty((V, 4, 3, 3, 2))
    _check(lib().nq_host_gauge_random_lx(*dims(L), seed, eps, _dp(conf)))
    return conf


def link_algebra(seed, lx, mu, eps):
    H = np.empty((3, 3, 2))
    U = np.empty((3, 3, 2))
    _check(lib().nq_host_link_algebra(seed, lx, mu, eps, _dp(H), _dp(U)))
    return H, U


def gauge_phase_lx(L, conf, theta_in_pi):
    th = np.ascontiguousarray(theta_in_pi, dtype=np.float64)
    _check(lib().nq_host_gauge_phase_lx(*dims(L), _dp(th), _dp(conf)))
    return conf


def read_ildg_gauge_conf(path, L):
    """nq_read_ildg_gauge_conf: (lx conf (V,4,3,3,2) in nissa's order, checksums
    (suma, sumb) read from the file and (suma, sumb) computed from its data)."""
    conf = np.empty((int(np.prod(L)), 4, 3, 3, 2))
    chk = (C.c_uint32 * 4)()
    _check(lib().nq_read_ildg_gauge_conf(str(path).encode(), *dims(L), _dp(conf), chk))
    return conf, (chk[0], chk[1]), (chk[2], chk[3])


def read_lime_real_vector(path, L, record="scidac-binary-data", nreals=24):
    """nq_read_lime_real_vector: (lx array (V, nreals), checksums read, checksums computed)."""
    out = np.empty((int(np.prod(L)), nreals))
    chk = (C.c_uint32 * 4)()
    _check(lib().nq_read_lime_real_vector(str(path).encode(), record.encode(), nreals, *dims(L), _dp(out), chk))
    return out, (chk[0], chk[1]), (chk[2], chk[3])


def unitarity_check_lx(L, conf):
    """nq_host_unitarity_check_lx: (average, maximum, number above 1e-13) of |U U^+ - 1|."""
    conf = np.ascontiguousarray(conf, dtype=np.float64)
    if conf.size != volume(L) * 72:
        raise NqError(NQ_EINVAL, f"conf of {conf.size} doubles, {volume(L) * 72} expected for {dims(L)}")
    out = (C.c_double * 2)()
    nb = C.c_int64(0)
    _check(lib().nq_host_unitarity_check_lx(*dims(L), _dp(conf),
                                            C.cast(out, _DP), C.byref(nb)))
    return out[0], out[1], nb.value


def conf_fingerprint(a):
    """nq_host_conf_fingerprint of a float64 host array."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = C.c_uint64(0)
    _check(lib().nq_host_conf_fingerprint(_dp(a), a.size, C.byref(out)))
    return out.value


def gauge_matches_lx(g, conf_lx):
    """True when the host lx conf still matches the one the handle was made from."""
    same = C.c_int(0)
    _check(lib().nq_gauge_matches_lx(g.h, _dp(np.ascontiguousarray(conf_lx, dtype=np.float64)), C.byref(same)))
    return bool(same.value)


def gauge_plaquette(ctx, g):
    d = C.c_double(0)
    _check(lib().nq_gauge_plaquette(ctx.h, g.h, C.byref(d)))
    return d.value


def gauge_stagphases_lx(L, conf):
    _check(lib().nq_host_gauge_stagphases_lx(*dims(L), _dp(conf)))
    return conf


def source_lx(L, seed, kind="gauss", ncomp=12, parity_mask=3, site0=0, sc=0):
    kinds = {"z2": 0, "gauss": 1, "point": 2}
    V = volume(L)
    out = np.empty((V, ncomp, 2)) if ncomp == 3 else np.empty((V, 4, 3, 2))
    _check(lib().nq_host_source_lx(*dims(L), seed, kinds[kind], ncomp, parity_mask, site0, sc, _dp(out)))
    return out


def split_lx(L, lx):
    V = volume(L)
    nd = lx.size // V
    e = np.empty((V // 2,) + lx.shape[1:])
    o = np.empty_like(e)
    _check(lib().nq_host_split_lx(*dims(L), nd, _dp(np.ascontiguousarray(lx)), _dp(e), _dp(o)))
    return e, o


def paste_lx(L, e, o):
    V = volume(L)
    lx = np.empty((V,) + e.shape[1:])
    nd = lx.size // V
    _check(lib().nq_host_paste_lx(*dims(L), nd, _dp(lx), _dp(np.ascontiguousarray(e)), _dp(np.ascontiguousarray(o))))
    return lx


# ----------------------------------------------------------------- device objects
class Ctx:
    def __init__(self, L, device=0):
        self.L = dims(L)
        self.V = volume(L)
        self.Vh = self.V // 2
        self.h = None
        # native handles of the device objects made in this context, with their
        # destroy functions.  Only handles are kept, never the objects: when the
        # cyclic garbage collector finalises a context together with its objects it
        # clears weak references first and runs the finalisers in any order, so the
        # context must be able to destroy whatever is still open by itself, and an
        # object finalised after its context must not touch the freed context.
        self._open = {}
        h = _P()
        _check(lib().nq_ctx_create(*self.L, device, C.byref(h)))
        self.h = h

    def _adopt(self, obj, destroy):
        key = id(obj)
        # the weak reference only serves to mark a live object closed; when it is
        # already cleared (the object is being collected) the handle alone suffices
        self._open[key] = (destroy, obj.h, weakref.ref(obj))
        obj._release = lambda: self._open.pop(key, None) is not None

    def close(self):
        if self.h:
            # device objects of this context must go before the context itself; a
            # live object is marked closed, so that using it afterwards passes a null
            # handle (NQ_EINVAL) and never a freed one
            for destroy, h, ref in reversed(list(self._open.values())):
                destroy(h)
                o = ref()
                if o is not None:
                    o.h = None
            self._open.clear()
            lib().nq_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().nq_ctx_sync(self.h))

    def site_tiles(self):
        """(t tile, x patch, y patch) of the device site order."""
        tb, tx, ty = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(lib().nq_ctx_site_tiles(self.h, C.byref(tb), C.byref(tx), C.byref(ty)))
        return tb.value, tx.value, ty.value

    def profiling(self, on=True):
        _check(lib().nq_ctx_set_profiling(self.h, 1 if on else 0))

    def reset_kernel_stats(self):
        _check(lib().nq_ctx_reset_kernel_stats(self.h))

    def kernel_stats(self):
        n = C.c_int(0)
        _check(lib().nq_ctx_kernel_count(self.h, C.byref(n)))
        out = {}
        buf = C.create_string_buffer(128)
        for i in range(n.value):
            la = C.c_int64(0)
            ms = C.c_double(0)
            by = C.c_double(0)
            _check(lib().nq_ctx_kernel_stat(self.h, i, buf, 128, C.byref(la), C.byref(ms), C.byref(by)))
            out[buf.value.decode()] = {"launches": la.value, "total_ms": ms.value, "bytes_per_launch": by.value}
        return out

    def gauge_lx(self, conf_lx):
        return Gauge(self, lx=conf_lx)

    def gauge_eo(self, evn, odd):
        return Gauge(self, evn=evn, odd=odd)

    def field(self, kind=SPINCOLOR, coverage=ODD, data=None):
        f = Field(self, kind, coverage)
        if data is not None:
            f.upload(data)
        return f


class Gauge:
    def __init__(self, ctx, lx=None, evn=None, odd=None):
        self.ctx = ctx
        self.h = None
        h = _P()
        if lx is not None:
            lx = np.ascontiguousarray(lx, dtype=np.float64)
            if lx.size != ctx.V * 72:
                raise NqError(NQ_EINVAL, f"lx conf of {lx.size} doubles, {ctx.V * 72} expected")
            _check(lib().nq_gauge_create_lx(ctx.h, _dp(lx), C.byref(h)))
        else:
            evn = np.ascontiguousarray(evn, dtype=np.float64)
            odd = np.ascontiguousarray(odd, dtype=np.float64)
            if evn.size != ctx.Vh * 72 or odd.size != ctx.Vh * 72:
                raise NqError(NQ_EINVAL, f"eo conf of {evn.size}+{odd.size} doubles, 2x{ctx.Vh * 72} expected")
            _check(lib().nq_gauge_create_eo(ctx.h, _dp(evn), _dp(odd), C.byref(h)))
        self.h = h
        ctx._adopt(self, lib().nq_gauge_destroy)

    def device_bytes(self):
        return lib().nq_gauge_device_bytes(self.h)

    def close(self):
        if self.h:
            if self._release():  # still registered: the context has not destroyed it
                lib().nq_gauge_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Clover:
    """Clover term of a gauge handle (nq_clover_*): clover_term(Cl, cSW, conf) on the device."""

    def __init__(self, ctx, g, csw):
        self.ctx = ctx
        self.csw = csw
        self.h = None
        h = _P()
        _check(lib().nq_clover_create(ctx.h, g.h, csw, C.byref(h)))
        self.h = h
        ctx._adopt(self, lib().nq_clover_destroy)

    def download_lx(self):
        out = np.empty((self.ctx.V, 4, 3, 3, 2))
        _check(lib().nq_clover_download_lx(self.h, _dp(out)))
        return out

    def close(self):
        if self.h:
            if self._release():  # still registered: the context has not destroyed it
                lib().nq_clover_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    def __init__(self, ctx, kind, coverage):
        self.ctx, self.kind, self.coverage = ctx, kind, coverage
        self.h = None
        h = _P()
        _check(lib().nq_field_create(ctx.h, kind, coverage, C.byref(h)))
        self.h = h
        ctx._adopt(self, lib().nq_field_destroy)

    @property
    def shape(self):
        n = self.ctx.V if self.coverage == LX else self.ctx.Vh
        return (n, 4, 3, 2) if self.kind == SPINCOLOR else (n, 3, 2)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == int(np.prod(self.shape)), (a.shape, self.shape)
        _check(lib().nq_field_upload(self.h, _dp(a)))

    def download(self):
        a = np.empty(self.shape)
        _check(lib().nq_field_download(self.h, _dp(a)))
        return a

    def norm2(self):
        d = C.c_double(0)
        _check(lib().nq_field_norm2(self.h, C.byref(d)))
        return d.value

    def close(self):
        if self.h:
            if self._release():  # still registered: the context has not destroyed it
                lib().nq_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- operators and solvers
def tm_hop(ctx, g, out, inp, anis=None):
    _check(lib().nq_tm_hop(ctx.h, g.h, _anis(anis), out.h, inp.h))


def tm_Dee(ctx, kappa, mu, out, inp, inverse=False):
    _check(lib().nq_tm_Dee(ctx.h, kappa, mu, 1 if inverse else 0, out.h, inp.h))


def tm_kern(ctx, g, kappa, mu, out, inp, anis=None):
    _check(lib().nq_tm_kern(ctx.h, g.h, kappa, _anis(anis), mu, out.h, inp.h))


def tm_kern_square(ctx, g, kappa, mu, out, inp, anis=None):
    _check(lib().nq_tm_kern_square(ctx.h, g.h, kappa, _anis(anis), mu, out.h, inp.h))


def tm_kern_square_noalloc(ctx, g, kappa, mu, out, inp, anis=None):
    """nq_tm_kern_square_noalloc: K K^+ with the library's static device scratch (no allocation)."""
    _check(lib().nq_tm_kern_square_noalloc(ctx.h, g.h, kappa, _anis(anis), mu, out.h, inp.h))


def tm_apply_Q(ctx, g, kappa, mu, out, inp, anis=None):
    _check(lib().nq_tm_apply_Q(ctx.h, g.h, kappa, _anis(anis), mu, out.h, inp.h))


def tm_apply_Q2(ctx, g, kappa, mu, out, inp):
    _check(lib().nq_tm_apply_Q2(ctx.h, g.h, kappa, mu, out.h, inp.h))


def tmclov_Dee(ctx, cl, kappa, mu, out, inp, dag=False, inverse=False):
    f = lib().nq_tmclov_inv_Dee if inverse else lib().nq_tmclov_Dee
    _check(f(ctx.h, cl.h, kappa, mu, int(dag), out.h, inp.h))


def tmclov_kern(ctx, g, cl, kappa, mu, out, inp, dag=False, anis=None):
    _check(lib().nq_tmclov_kern(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, int(dag), out.h, inp.h))


def tmclov_kern_square(ctx, g, cl, kappa, mu, out, inp, anis=None):
    _check(lib().nq_tmclov_kern_square(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, out.h, inp.h))


def tmclov_apply_Q(ctx, g, cl, kappa, mu, out, inp, anis=None):
    _check(lib().nq_tmclov_apply_Q(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, out.h, inp.h))


def tmclov_apply_Q2(ctx, g, cl, kappa, mu, out, inp, anis=None):
    _check(lib().nq_tmclov_apply_Q2(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, out.h, inp.h))


def tmclovQ2_cg(ctx, g, cl, kappa, mu, niter, residue, sol, src, guess=None, anis=None):
    st = SolveStats()
    rc = lib().nq_tmclovQ2_cg(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, niter, residue, sol.h,
                              guess.h if guess is not None else None, src.h, C.byref(st))
    return _solve(rc, st)


def tmclovQ2_cgm(ctx, g, cl, kappa, masses, niter, residues, sols, src, anis=None):
    """inv_tmclovQ2_cgm: sols[k] per mass; returns stats and the FP64 true residual per mass."""
    n = len(masses)
    ma = (C.c_double * n)(*masses)
    rs = (C.c_double * n)(*residues)
    so = (_P * n)(*[x.h for x in sols])
    st = SolveStats()
    tr = (C.c_double * n)()
    rc = lib().nq_tmclovQ2_cgm(ctx.h, g.h, cl.h, kappa, _anis(anis), n, ma, niter, rs, so, src.h, C.byref(st), tr)
    if rc not in (NQ_OK, NQ_ENOCONV):
        _check(rc)
    return dict(st.as_dict(), rc=rc, true_res_sq=list(tr))


def _solve(rc, st):
    if rc not in (NQ_OK, NQ_ENOCONV):
        _check(rc)
    d = st.as_dict()
    d["rc"] = rc
    return d


def tm_cg_kernsq(ctx, g, kappa, mu, niter, residue, sol, src, prec="f64", guess=None, anis=None):
    st = SolveStats()
    rc = lib().nq_tm_cg_kernsq(ctx.h, g.h, kappa, _anis(anis), mu, niter, residue, PREC[prec], sol.h,
                               guess.h if guess is not None else None, src.h, C.byref(st))
    return _solve(rc, st)


def tm_solve_eoprec(ctx, g, kappa, mu, niter, residue, sol, src, prec="f64", anis=None, guess=None):
    """inv_tmD_cg_eoprec; guess (odd field, optional) starts the eq. 9 CG and receives its solution."""
    st = SolveStats()
    chk = C.c_double(0)
    rc = lib().nq_tm_solve_eoprec(ctx.h, g.h, kappa, _anis(anis), mu, niter, residue, PREC[prec], sol.h,
                                  guess.h if guess is not None else None, src.h, C.byref(st), C.byref(chk))
    d = _solve(rc, st)
    d["check_rel_res_sq"] = chk.value
    return d


def tm_solve_eoprec_host(ctx, g, kappa, mu, niter, residue, src_lx, prec="f64", anis=None, guess=None):
    """nq_tm_solve_eoprec_host on nissa host buffers: returns (solution lx, stats); guess
    (odd host array (Vh,4,3,2), optional) is updated in place with the eq. 9 solution."""
    src_lx = np.ascontiguousarray(src_lx, dtype=np.float64)
    if src_lx.size != ctx.V * 24:
        raise NqError(NQ_EINVAL, f"source of {src_lx.size} doubles, {ctx.V * 24} expected")
    if guess is not None and (guess.size != ctx.Vh * 24 or guess.dtype != np.float64 or not guess.flags.c_contiguous):
        raise NqError(NQ_EINVAL, "guess must be a contiguous float64 odd field")
    sol = np.empty((ctx.V, 4, 3, 2))
    st = SolveStats()
    chk = C.c_double(0)
    rc = lib().nq_tm_solve_eoprec_host(ctx.h, g.h, kappa, _anis(anis), mu, niter, residue, PREC[prec], _dp(sol),
                                       _dp(guess), _dp(src_lx), C.byref(st), C.byref(chk))
    d = _solve(rc, st)
    d["check_rel_res_sq"] = chk.value
    return sol, d


def tm_kern_square_host(ctx, g, kappa, mu, inp_odd, anis=None):
    """nq_tm_kern_square_host: K K^+ on an odd host array (Vh,4,3,2)."""
    inp_odd = np.ascontiguousarray(inp_odd, dtype=np.float64)
    if inp_odd.size != ctx.Vh * 24:
        raise NqError(NQ_EINVAL, f"input of {inp_odd.size} doubles, {ctx.Vh * 24} expected")
    out = np.empty((ctx.Vh, 4, 3, 2))
    _check(lib().nq_tm_kern_square_host(ctx.h, g.h, kappa, _anis(anis), mu, _dp(out), _dp(inp_odd)))
    return out


def tmclov_cg_kernsq(ctx, g, cl, kappa, mu, niter, residue, sol, src, guess=None, anis=None, prec="f64"):
    st = SolveStats()
    rc = lib().nq_tmclov_cg_kernsq(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, niter, residue, PREC[prec], sol.h,
                                   guess.h if guess is not None else None, src.h, C.byref(st))
    return _solve(rc, st)


def tmclov_solve_eoprec(ctx, g, cl, kappa, mu, niter, residue, sol, src, anis=None, prec="f64"):
    st = SolveStats()
    chk = C.c_double(0)
    rc = lib().nq_tmclov_solve_eoprec(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, niter, residue, PREC[prec], sol.h,
                                      src.h, C.byref(st), C.byref(chk))
    d = _solve(rc, st)
    d["check_rel_res_sq"] = chk.value
    return d


def tmclov_solve_eoprec_multi(ctx, g, cl, kappa, mu, niter, residue, sols, srcs, prec="f16", anis=None):
    """nq_tmclov_solve_eoprec_multi: the twisted clover eo solve of several sources at one mu."""
    n = len(srcs)
    assert len(sols) == n
    so = (_P * n)(*[s.h for s in sols])
    sr = (_P * n)(*[s.h for s in srcs])
    st = (SolveStats * n)()
    rc = lib().nq_tmclov_solve_eoprec_multi(ctx.h, g.h, cl.h, kappa, _anis(anis), mu, n, niter, residue, PREC[prec],
                                            so, sr, st)
    if rc not in (NQ_OK, NQ_ENOCONV):
        _check(rc)
    return [dict(st[i].as_dict(), rc=rc) for i in range(n)]


def tmQ2_cg(ctx, g, kappa, mu, niter, residue, sol, src, guess=None):
    st = SolveStats()
    rc = lib().nq_tmQ2_cg(ctx.h, g.h, kappa, mu, niter, residue, sol.h, guess.h if guess is not None else None,
                          src.h, C.byref(st))
    return _solve(rc, st)


def tm_solve_eoprec_multi(ctx, g, kappa, mus, niter, residue, sols, srcs, prec="f16", anis=None, guesses=None):
    n = len(srcs)
    assert len(sols) == n and len(mus) == n
    mu = (C.c_double * n)(*mus)
    so = (_P * n)(*[s.h for s in sols])
    sr = (_P * n)(*[s.h for s in srcs])
    gs = None
    if guesses is not None:
        assert len(guesses) == n
        gs = (_P * n)(*[x.h if x is not None else None for x in guesses])
    st = (SolveStats * n)()
    rc = lib().nq_tm_solve_eoprec_multi(ctx.h, g.h, kappa, _anis(anis), n, mu, niter, residue, PREC[prec], so, gs, sr,
                                        st)
    if rc not in (NQ_OK, NQ_ENOCONV):
        _check(rc)
    return [dict(st[i].as_dict(), rc=rc) for i in range(n)]


def st_2Doe(ctx, g, out, inp):
    _check(lib().nq_st_2Doe(ctx.h, g.h, out.h, inp.h))


def st_Deo_half(ctx, g, out, inp):
    _check(lib().nq_st_Deo_half(ctx.h, g.h, out.h, inp.h))


def st_DDee(ctx, g, mass2, out, inp):
    _check(lib().nq_st_DDee(ctx.h, g.h, mass2, out.h, inp.h))


def st_cg(ctx, g, mass2, niter, residue, sol, src, guess=None):
    st = SolveStats()
    rc = lib().nq_st_cg(ctx.h, g.h, mass2, niter, residue, sol.h, guess.h if guess is not None else None, src.h,
                        C.byref(st))
    return _solve(rc, st)


def st_cgm(ctx, g, shifts, niter, residues, sols, srcs, prec="f64"):
    """sols[r][k] for source r and shift k."""
    nr, ns = len(srcs), len(shifts)
    sh = (C.c_double * ns)(*shifts)
    rs = (C.c_double * ns)(*residues)
    so = (_P * (nr * ns))(*[sols[r][k].h for r in range(nr) for k in range(ns)])
    sr = (_P * nr)(*[s.h for s in srcs])
    st = (SolveStats * nr)()
    tr = (C.c_double * (nr * ns))()
    rc = lib().nq_st_cgm(ctx.h, g.h, nr, ns, sh, niter, rs, PREC[prec], so, sr, st, tr)
    if rc not in (NQ_OK, NQ_ENOCONV):
        _check(rc)
    return [dict(st[r].as_dict(), rc=rc, true_res_sq=[tr[r * ns + k] for k in range(ns)]) for r in range(nr)]


def st_apply_D(ctx, g, m, out_e, out_o, in_e, in_o, dag=False):
    _check(lib().nq_st_apply_D(ctx.h, g.h, m, int(dag), out_e.h if out_e is not None else None,
                               out_o.h if out_o is not None else None, in_e.h, in_o.h))


def st_solve_D(ctx, g, m, niter, residue, sol_e, sol_o, src_e, src_o, guess=None):
    st = SolveStats()
    chk = C.c_double(0)
    rc = lib().nq_st_solve_D(ctx.h, g.h, m, niter, residue, sol_e.h, sol_o.h if sol_o is not None else None,
                             guess.h if guess is not None else None, src_e.h, src_o.h, C.byref(st), C.byref(chk))
    out = _solve(rc, st)
    out["check_rel_res_sq"] = chk.value
    return out


def st_rootst_force(ctx, g, u1b_e, u1b_o, poles, weights, niter, residue, pf, Fe=None, Fo=None, prec="f64"):
    """summ_the_rootst_eoimpr_quark_force: F (host, (Vh,4,3,3,2) per parity) summed into."""
    Vh = ctx.Vh
    Fe = np.zeros((Vh, 4, 3, 3, 2)) if Fe is None else Fe
    Fo = np.zeros((Vh, 4, 3, 3, 2)) if Fo is None else Fo
    n = len(poles)
    st = SolveStats()
    rc = lib().nq_st_rootst_force(ctx.h, g.h, _dp(np.ascontiguousarray(u1b_e)), _dp(np.ascontiguousarray(u1b_o)), n,
                                  (C.c_double * n)(*poles), (C.c_double * n)(*weights), niter, residue, PREC[prec],
                                  pf.h, _dp(Fe), _dp(Fo), C.byref(st))
    out = _solve(rc, st)
    return Fe, Fo, out


def tmQ2_cgm(ctx, g, kappa, masses, niter, residues, sols, src, g5=False):
    """inv_tmQ2_cgm (or inv_tmDQ_cgm with g5=True): sols[k] per mass; returns stats and
    the FP64 true residual per mass."""
    n = len(masses)
    ma = (C.c_double * n)(*masses)
    rs = (C.c_double * n)(*residues)
    so = (_P * n)(*[x.h for x in sols])
    st = SolveStats()
    tr = (C.c_double * n)()
    fn = lib().nq_tmDQ_cgm if g5 else lib().nq_tmQ2_cgm
    rc = fn(ctx.h, g.h, kappa, n, ma, niter, rs, so, src.h, C.byref(st), tr)
    if rc not in (NQ_OK, NQ_ENOCONV):
        _check(rc)
    return dict(st.as_dict(), rc=rc, true_res_sq=list(tr))


def st_cgm_rat_sum(ctx, g, cons, poles, weights, niter, residue, chis, srcs, prec="f64"):
    """chi_r = cons src_r + sum_k weights[k] (D2ee + poles[k])^-1 src_r for every source."""
    nr, d = len(srcs), len(poles)
    assert len(weights) == d and len(chis) == nr
    po = (C.c_double * d)(*poles)
    we = (C.c_double * d)(*weights)
    ch = (_P * nr)(*[x.h for x in chis])
    sr = (_P * nr)(*[x.h for x in srcs])
    st = (SolveStats * nr)()
    rc = lib().nq_st_cgm_rat_sum(ctx.h, g.h, nr, d, cons, po, we, niter, residue, PREC[prec], ch, sr, st)
    if rc not in (NQ_OK, NQ_ENOCONV):
        _check(rc)
    return [dict(st[r].as_dict(), rc=rc) for r in range(nr)]


def bench_tm_hop_multi(ctx, g, nrhs, napply, inp, prec="f16"):
    ms = C.c_double(0)
    by = C.c_double(0)
    _check(lib().nq_bench_tm_hop_multi(ctx.h, g.h, PREC[prec], nrhs, napply, inp.h, C.byref(ms), C.byref(by)))
    return ms.value, by.value


def bench_tm_hop(ctx, g, napply, out, inp, prec="f64"):
    ms = C.c_double(0)
    by = C.c_double(0)
    _check(lib().nq_bench_tm_hop(ctx.h, g.h, PREC[prec], napply, out.h, inp.h, C.byref(ms), C.byref(by)))
    return ms.value, by.value


class Server:
    """Propagator-solve server (nq_server_*): resident gauge fields per volume,
    continuous batching of the sources of queued requests."""

    def __init__(self, device=0, hbm_budget=0, max_groups=1):
        h = _P()
        _check(lib().nq_server_create(device, int(hbm_budget), max_groups, C.byref(h)))
        self.h = h
        self._keep = {}

    def add_volume(self, L, conf_lx):
        conf = np.ascontiguousarray(conf_lx, dtype=np.float64)
        vid = C.c_int()
        _check(lib().nq_server_add_volume(self.h, *dims(L), _dp(conf), C.byref(vid)))
        return vid.value

    def submit(self, volume, kappa, mu, residue, prec="f16", niter=100000, point_site=None, nsrc=12,
               srcs=None, sols=None, csw=0.0):
        """Point sources (spin-colour 0..nsrc-1 at lx site point_site) or host sources
        `srcs` (list of lx arrays); `sols` (list of lx arrays) receive the solutions;
        csw > 0 makes it a twisted clover request."""
        r = ServeRequest()
        r.volume, r.prec, r.kappa, r.mu, r.residue, r.niter = volume, PREC[prec], kappa, mu, residue, niter
        r.csw = csw
        keep = []
        if srcs is not None:
            srcs = [np.ascontiguousarray(x, dtype=np.float64) for x in srcs]
            r.nsrc, r.src_kind = len(srcs), 0
            arr = (_DP * len(srcs))(*[_dp(x) for x in srcs])
            r.src = C.cast(arr, C.c_void_p)
            keep += [srcs, arr]
        else:
            r.nsrc, r.src_kind, r.site0 = nsrc, 1, int(point_site or 0)
        if sols is not None:
            assert len(sols) == r.nsrc and all(x.flags.c_contiguous and x.dtype == np.float64 for x in sols)
            arr = (_DP * len(sols))(*[_dp(x) for x in sols])
            r.sol = C.cast(arr, C.c_void_p)
            keep += [sols, arr]
        rid = C.c_int64()
        _check(lib().nq_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = keep
        return rid.value

    def wait(self, rid, timeout=600.0):
        res = ServeResult()
        _check(lib().nq_server_wait(self.h, rid, timeout, C.byref(res)))
        self._keep.pop(rid, None)
        return res.as_dict()

    def clock(self):
        return lib().nq_server_clock(self.h)

    def device_bytes(self):
        return lib().nq_server_device_bytes(self.h)

    def close(self):
        if self.h:
            lib().nq_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
