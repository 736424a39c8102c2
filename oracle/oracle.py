"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_DP = C.POINTER(C.c_double)


class CgStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("rel_res_sq", C.c_double), ("converged", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", HERE], check=True, capture_output=True)
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_omp_max_threads": ([], _I),
            "or_geom_create": ([_I, _I, _I, _I], _P),
            "or_geom_destroy": ([_P], None),
            "or_geom_vol": ([_P], C.c_int64),
            "or_geom_lx_of_eo": ([_P, _I, C.c_int64], C.c_int64),
            "or_geom_eo_of_lx": ([_P, C.c_int64], C.c_int64),
            "or_geom_parity": ([_P, C.c_int64], _I),
            "or_geom_neigh_lx": ([_P, C.c_int64, _I, _I], C.c_int64),
            "or_split_lx": ([_P, _I, _DP, _DP, _DP], None),
            "or_paste_lx": ([_P, _I, _DP, _DP, _DP], None),
            "or_tm_hop": ([_P, _I, _DP, _DP, _DP, _DP, _DP], None),
            "or_tmDee": ([_P, _DP, _D, _D, _DP], None),
            "or_inv_tmDee": ([_P, _DP, _D, _D, _DP], None),
            "or_tm_put_together_g5": ([_P, _DP, _DP], None),
            "or_tmDkern": ([_P, _DP, _DP, _DP, _DP, _D, _DP, _D, _DP], None),
            "or_tmDkern_square": ([_P, _DP, _DP, _DP, _DP, _DP, _D, _DP, _D, _DP], None),
            "or_clover_term_lx": ([_P, _DP, _DP, _D], None),
            "or_rootst_force": ([_P, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _I, _DP, _DP, _I, _D], _I),
            "or_tmclovDee": ([_P, _DP, _DP, _D, _D, _I, _DP], None),
            "or_tmclov_invert": ([_P, _DP, _DP, _D, _D], None),
            "or_inv_tmclovDee": ([_P, _DP, _DP, _I, _DP], None),
            "or_tmclovDkern": ([_P, _DP, _DP, _DP, _DP, _D, _DP, _DP, _DP, _I, _D, _DP], None),
            "or_tmclovDkern_square": ([_P, _DP, _DP, _DP, _DP, _DP, _D, _DP, _DP, _DP, _D, _DP], None),
            "or_apply_tmclovQ": ([_P, _DP, _DP, _D, _DP, _DP, _D, _DP], None),
            "or_tmclov_solve_eoprec": ([_P, _DP, _DP, _D, _DP, _D, _D, _I, _D, _DP, C.POINTER(CgStats), _DP], _I),
            "or_apply_tmQ": ([_P, _DP, _DP, _D, _DP, _D, _DP], None),
            "or_apply_tmQ2": ([_P, _DP, _DP, _D, _DP, _D, _DP], None),
            "or_st_2Doe": ([_P, _DP, _DP, _DP, _DP], None),
            "or_st_Deo_half": ([_P, _DP, _DP, _DP, _DP], None),
            "or_stD2ee_m2": ([_P, _DP, _DP, _DP, _DP, _D, _DP], None),
            "or_apply_stD": ([_P, _DP, _DP, _DP, _DP, _D, _DP, _DP], None),
            "or_scalar_prod": ([C.c_int64, _I, _DP, _DP], _D),
            "or_tm_cg_kernsq": ([_P, _DP, _DP, _DP, _DP, _D, _DP, _D, _I, _D, _DP, C.POINTER(CgStats)], _I),
            "or_tm_solve_eoprec": ([_P, _DP, _DP, _D, _DP, _D, _I, _D, _DP, C.POINTER(CgStats), _DP], _I),
            "or_tmQ2_cg": ([_P, _DP, _DP, _DP, _D, _D, _I, _D, _DP, C.POINTER(CgStats)], _I),
            "or_st_cg": ([_P, _DP, _DP, _DP, _D, _I, _D, _DP, C.POINTER(CgStats)], _I),
            "or_st_cgm": ([_P, C.POINTER(_DP), _DP, _DP, _I, _DP, _I, _DP, _DP, C.POINTER(CgStats), _DP], _I),
            "or_tmQ2_cgm": ([_P, C.POINTER(_DP), _DP, _D, _I, _DP, _I, _DP, _DP, C.POINTER(CgStats), _DP], _I),
            "or_st_rat_summ": ([_P, _DP, _DP, _DP, _I, _D, _DP, _DP, _I, _D, _DP, C.POINTER(CgStats)], _I),
            "or_tm_cg128_kernsq": ([_P, _DP, _DP, _DP, _D, _DP, _D, _D, _DP, _DP, C.POINTER(C.c_int)], _I),
            "or_tmDkern_square_q": ([_P, _DP, _DP, _DP, _D, _DP, _D, _DP], None),
            "or_plaquette_lx": ([_P, _DP], _D),
            "or_gauge_transform_lx": ([_P, _DP, _DP], None),
            "or_color_rotate_lx": ([_P, _I, _DP, _DP], None),
        }
        for n, (a, r) in sig.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_DP)


def _anis(anis):
    return None if anis is None else C.pointer(C.c_double(anis))


class Geom:
    """Lattice geometry T x LX x LY x LZ (extents even)."""

    def __init__(self, L):
        self.L = tuple(int(x) for x in L)
        self.h = lib().or_geom_create(*self.L)
        if not self.h:
            raise ValueError(f"bad lattice {L}")
        self.V = int(np.prod(self.L))
        self.Vh = self.V // 2

    def __del__(self):
        try:
            if self.h:
                lib().or_geom_destroy(self.h)
        except Exception:
            pass

    def lx_of_eo(self, par, ieo):
        return lib().or_geom_lx_of_eo(self.h, par, ieo)

    def eo_of_lx(self, lx):
        return lib().or_geom_eo_of_lx(self.h, lx)

    def parity(self, lx):
        return lib().or_geom_parity(self.h, lx)

    def neigh(self, lx, mu, up):
        return lib().or_geom_neigh_lx(self.h, lx, mu, 1 if up else 0)

    def split(self, lx):
        nd = lx.size // self.V
        e = np.empty((self.Vh,) + lx.shape[1:])
        o = np.empty_like(e)
        lib().or_split_lx(self.h, nd, _dp(np.ascontiguousarray(lx)), _dp(e), _dp(o))
        return e, o

    def paste(self, e, o):
        lx = np.empty((self.V,) + e.shape[1:])
        nd = lx.size // self.V
        lib().or_paste_lx(self.h, nd, _dp(lx), _dp(np.ascontiguousarray(e)), _dp(np.ascontiguousarray(o)))
        return lx

    # -- tm
    def tm_hop(self, out_par, ce, co, inp, anis=None):
        out = np.empty((self.Vh, 4, 3, 2))
        lib().or_tm_hop(self.h, out_par, _dp(out), _dp(ce), _dp(co), _anis(anis), _dp(np.ascontiguousarray(inp)))
        return out

    def tmDee(self, kappa, mu, inp, inverse=False):
        out = np.empty((self.Vh, 4, 3, 2))
        f = lib().or_inv_tmDee if inverse else lib().or_tmDee
        f(self.h, _dp(out), kappa, mu, _dp(np.ascontiguousarray(inp)))
        return out

    def tmDkern(self, ce, co, kappa, mu, inp, anis=None):
        out = np.empty((self.Vh, 4, 3, 2))
        t = np.empty_like(out)
        lib().or_tmDkern(self.h, _dp(out), _dp(t), _dp(ce), _dp(co), kappa, _anis(anis), mu,
                         _dp(np.ascontiguousarray(inp)))
        return out

    def tmDkern_square(self, ce, co, kappa, mu, inp, anis=None):
        out = np.empty((self.Vh, 4, 3, 2))
        t1 = np.empty_like(out)
        t2 = np.empty_like(out)
        lib().or_tmDkern_square(self.h, _dp(out), _dp(t1), _dp(t2), _dp(ce), _dp(co), kappa, _anis(anis), mu,
                                _dp(np.ascontiguousarray(inp)))
        return out

    def tmDkern_square_q(self, ce, co, kappa, mu, inp, anis=None):
        out = np.empty((self.Vh, 4, 3, 2))
        lib().or_tmDkern_square_q(self.h, _dp(out), _dp(ce), _dp(co), kappa, _anis(anis), mu,
                                  _dp(np.ascontiguousarray(inp)))
        return out

    # -- twisted clover
    def clover_term(self, conf_lx, csw):
        """clover_term(Cl, cSW, conf): [V][4][3][3][2] (G, H, I, J)."""
        cl = np.empty((self.V, 4, 3, 3, 2))
        lib().or_clover_term_lx(self.h, _dp(cl), _dp(np.ascontiguousarray(conf_lx)), csw)
        return cl

    def tmclovDee(self, cl_par, kappa, mu, inp, dag=False):
        out = np.empty((self.Vh, 4, 3, 2))
        lib().or_tmclovDee(self.h, _dp(out), _dp(np.ascontiguousarray(cl_par)), kappa, mu, int(dag),
                           _dp(np.ascontiguousarray(inp)))
        return out

    def tmclov_invert(self, cl_par, kappa, mu):
        inv = np.empty((self.Vh, 2, 6, 6, 2))
        lib().or_tmclov_invert(self.h, _dp(inv), _dp(np.ascontiguousarray(cl_par)), kappa, mu)
        return inv

    def inv_tmclovDee(self, inv, inp, dag=False):
        out = np.empty((self.Vh, 4, 3, 2))
        lib().or_inv_tmclovDee(self.h, _dp(out), _dp(np.ascontiguousarray(inv)), int(dag),
                               _dp(np.ascontiguousarray(inp)))
        return out

    def tmclovDkern(self, ce, co, kappa, cl_odd, inv_evn, mu, inp, dag=False, anis=None):
        out = np.empty((self.Vh, 4, 3, 2))
        t = np.empty_like(out)
        lib().or_tmclovDkern(self.h, _dp(out), _dp(t), _dp(ce), _dp(co), kappa, _anis(anis),
                             _dp(np.ascontiguousarray(cl_odd)), _dp(np.ascontiguousarray(inv_evn)), int(dag), mu,
                             _dp(np.ascontiguousarray(inp)))
        return out

    def tmclovDkern_square(self, ce, co, kappa, cl_odd, inv_evn, mu, inp, anis=None):
        out = np.empty((self.Vh, 4, 3, 2))
        t1 = np.empty_like(out)
        t2 = np.empty_like(out)
        lib().or_tmclovDkern_square(self.h, _dp(out), _dp(t1), _dp(t2), _dp(ce), _dp(co), kappa, _anis(anis),
                                    _dp(np.ascontiguousarray(cl_odd)), _dp(np.ascontiguousarray(inv_evn)), mu,
                                    _dp(np.ascontiguousarray(inp)))
        return out

    def apply_tmclovQ(self, conf_lx, kappa, cl_lx, mu, inp, anis=None):
        out = np.empty((self.V, 4, 3, 2))
        lib().or_apply_tmclovQ(self.h, _dp(out), _dp(np.ascontiguousarray(conf_lx)), kappa, _anis(anis),
                               _dp(np.ascontiguousarray(cl_lx)), mu, _dp(np.ascontiguousarray(inp)))
        return out

    def tmclov_solve_eoprec(self, conf_lx, kappa, csw, mu, niter, residue, src_lx, anis=None):
        sol = np.empty((self.V, 4, 3, 2))
        st = CgStats()
        chk = C.c_double(0)
        rc = lib().or_tmclov_solve_eoprec(self.h, _dp(sol), _dp(np.ascontiguousarray(conf_lx)), kappa, _anis(anis),
                                          csw, mu, niter, residue, _dp(np.ascontiguousarray(src_lx)), C.byref(st),
                                          C.byref(chk))
        return sol, {"rc": rc, "iters": st.iters, "rel_res_sq": st.rel_res_sq, "check": chk.value}

    # -- rooted staggered force
    def rootst_force(self, ce, co, u1e, u1o, pf, poles, weights, niter, residue, Fe=None, Fo=None):
        """summ_the_rootst_eoimpr_quark_force: ce/co pure links, u1e/u1o (Vh,4,2); F summed into."""
        Fe = np.zeros((self.Vh, 4, 3, 3, 2)) if Fe is None else Fe
        Fo = np.zeros((self.Vh, 4, 3, 3, 2)) if Fo is None else Fo
        n = len(poles)
        rc = lib().or_rootst_force(self.h, _dp(Fe), _dp(Fo), _dp(np.ascontiguousarray(ce)), _dp(np.ascontiguousarray(co)),
                                   _dp(np.ascontiguousarray(u1e)), _dp(np.ascontiguousarray(u1o)),
                                   _dp(np.ascontiguousarray(pf)), n, (C.c_double * n)(*poles),
                                   (C.c_double * n)(*weights), niter, residue)
        return Fe, Fo, rc

    def apply_tmQ(self, conf_lx, kappa, mu, inp, anis=None):
        out = np.empty((self.V, 4, 3, 2))
        lib().or_apply_tmQ(self.h, _dp(out), _dp(conf_lx), kappa, _anis(anis), mu, _dp(np.ascontiguousarray(inp)))
        return out

    def apply_tmQ2(self, conf_lx, kappa, mu, inp):
        out = np.empty((self.V, 4, 3, 2))
        lib().or_apply_tmQ2(self.h, _dp(out), _dp(conf_lx), kappa, None, mu, _dp(np.ascontiguousarray(inp)))
        return out

    def tm_cg_kernsq(self, ce, co, kappa, mu, niter, residue, src, anis=None, guess=None):
        sol = np.empty((self.Vh, 4, 3, 2))
        st = CgStats()
        rc = lib().or_tm_cg_kernsq(self.h, _dp(sol), _dp(guess), _dp(ce), _dp(co), kappa, _anis(anis), mu, niter,
                                   residue, _dp(np.ascontiguousarray(src)), C.byref(st))
        return sol, {"rc": rc, "iters": st.iters, "rel_res_sq": st.rel_res_sq, "converged": st.converged}

    def tm_solve_eoprec(self, conf_lx, kappa, mu, niter, residue, src_lx, anis=None):
        sol = np.empty((self.V, 4, 3, 2))
        st = CgStats()
        chk = C.c_double(0)
        rc = lib().or_tm_solve_eoprec(self.h, _dp(sol), _dp(conf_lx), kappa, _anis(anis), mu, niter, residue,
                                      _dp(np.ascontiguousarray(src_lx)), C.byref(st), C.byref(chk))
        return sol, {"rc": rc, "iters": st.iters, "rel_res_sq": st.rel_res_sq, "check": chk.value}

    def tmQ2_cg(self, conf_lx, kappa, mu, niter, residue, src):
        sol = np.empty((self.V, 4, 3, 2))
        st = CgStats()
        rc = lib().or_tmQ2_cg(self.h, _dp(sol), None, _dp(conf_lx), kappa, mu, niter, residue,
                              _dp(np.ascontiguousarray(src)), C.byref(st))
        return sol, {"rc": rc, "iters": st.iters, "rel_res_sq": st.rel_res_sq}

    def tmQ2_cgm(self, conf_lx, kappa, masses, niter, residues, src):
        n = len(masses)
        sols = [np.empty((self.V, 4, 3, 2)) for _ in range(n)]
        ptrs = (_DP * n)(*[_dp(s) for s in sols])
        st = CgStats()
        tr = (C.c_double * n)()
        rc = lib().or_tmQ2_cgm(self.h, ptrs, _dp(conf_lx), kappa, n, (C.c_double * n)(*masses), niter,
                               (C.c_double * n)(*residues), _dp(np.ascontiguousarray(src)), C.byref(st), tr)
        return sols, {"rc": rc, "iters": st.iters, "true_res_sq": list(tr)}

    def tm_cg128_kernsq(self, ce, co, kappa, mu, target, src, anis=None):
        sol = np.empty((self.Vh, 4, 3, 2))
        fr = C.c_double(0)
        outer = C.c_int(0)
        rc = lib().or_tm_cg128_kernsq(self.h, _dp(sol), _dp(ce), _dp(co), kappa, _anis(anis), mu, target,
                                      _dp(np.ascontiguousarray(src)), C.byref(fr), C.byref(outer))
        return sol, {"rc": rc, "rel_res_sq": fr.value, "outer": outer.value}

    # -- staggered
    def st_2Doe(self, ce, co, inp_e):
        out = np.empty((self.Vh, 3, 2))
        lib().or_st_2Doe(self.h, _dp(out), _dp(ce), _dp(co), _dp(np.ascontiguousarray(inp_e)))
        return out

    def st_Deo_half(self, ce, co, inp_o):
        out = np.empty((self.Vh, 3, 2))
        lib().or_st_Deo_half(self.h, _dp(out), _dp(ce), _dp(co), _dp(np.ascontiguousarray(inp_o)))
        return out

    def stD2ee_m2(self, ce, co, mass2, inp):
        out = np.empty((self.Vh, 3, 2))
        t = np.empty_like(out)
        lib().or_stD2ee_m2(self.h, _dp(out), _dp(ce), _dp(co), _dp(t), mass2, _dp(np.ascontiguousarray(inp)))
        return out

    def apply_stD(self, ce, co, m, ie, io):
        oe = np.empty((self.Vh, 3, 2))
        oo = np.empty_like(oe)
        lib().or_apply_stD(self.h, _dp(oe), _dp(oo), _dp(ce), _dp(co), m, _dp(np.ascontiguousarray(ie)),
                           _dp(np.ascontiguousarray(io)))
        return oe, oo

    def st_cg(self, ce, co, mass2, niter, residue, src):
        sol = np.empty((self.Vh, 3, 2))
        st = CgStats()
        rc = lib().or_st_cg(self.h, _dp(sol), _dp(ce), _dp(co), mass2, niter, residue,
                            _dp(np.ascontiguousarray(src)), C.byref(st))
        return sol, {"rc": rc, "iters": st.iters, "rel_res_sq": st.rel_res_sq}

    def st_cgm(self, ce, co, shifts, niter, residues, src):
        n = len(shifts)
        sols = [np.empty((self.Vh, 3, 2)) for _ in range(n)]
        ptrs = (_DP * n)(*[_dp(s) for s in sols])
        st = CgStats()
        tr = (C.c_double * n)()
        rc = lib().or_st_cgm(self.h, ptrs, _dp(ce), _dp(co), n, (C.c_double * n)(*shifts), niter,
                             (C.c_double * n)(*residues), _dp(np.ascontiguousarray(src)), C.byref(st), tr)
        return sols, {"rc": rc, "iters": st.iters, "true_res_sq": list(tr)}

    def st_rat_summ(self, ce, co, cons, poles, weights, niter, residue, src):
        """chi = cons src + sum_k weights[k] (A + poles[k])^-1 src, A = -1/4 Deo Doe (massless)."""
        n = len(poles)
        chi = np.empty((self.Vh, 3, 2))
        st = CgStats()
        rc = lib().or_st_rat_summ(self.h, _dp(chi), _dp(ce), _dp(co), n, cons, (C.c_double * n)(*poles),
                                  (C.c_double * n)(*weights), niter, residue, _dp(np.ascontiguousarray(src)),
                                  C.byref(st))
        return chi, {"rc": rc, "iters": st.iters}

    def plaquette(self, conf_lx):
        return lib().or_plaquette_lx(self.h, _dp(np.ascontiguousarray(conf_lx)))

    def gauge_transform(self, conf_lx, omega):
        c = np.ascontiguousarray(conf_lx).copy()
        lib().or_gauge_transform_lx(self.h, _dp(c), _dp(np.ascontiguousarray(omega)))
        return c

    def color_rotate(self, psi_lx, omega, nspin):
        p = np.ascontiguousarray(psi_lx).copy()
        lib().or_color_rotate_lx(self.h, nspin, _dp(p), _dp(np.ascontiguousarray(omega)))
        return p


def omp_max_threads():
    """Threads the oracle's OpenMP loops run on."""
    return lib().or_omp_max_threads()


def scalar_prod(a, b, nd):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return lib().or_scalar_prod(a.size // nd, nd, _dp(a), _dp(b))
