"""Shared test helpers: nissa's gamma matrices written out as data, complex
views of nissa-layout arrays, norms and standard synthetic inputs."""
from __future__ import annotations

import numpy as np

import nissa_amd as nq


def gamma_from_rows(rows):
    """rows: 4 tuples (column, re, im) -- row d of the matrix has entry re+i*im
    in column `column` (nissa's init_dirac convention, src/new_types/dirac.cpp:43-58,
    applied as out[d] = entr[d] * in[pos[d]], src/new_types/su3_op.hpp:1953-1959)."""
    g = np.zeros((4, 4), complex)
    for d, (c, re, im) in enumerate(rows):
        g[d, c] = re + 1j * im
    return g


# nissa's base_gamma table (src/new_types/dirac.cpp:45-50): [0]=1, [1..3]=gamma_{x,y,z},
# [4]=gamma_t, [5]=gamma_5
GAMMA = {
    0: gamma_from_rows([(0, 1, 0), (1, 1, 0), (2, 1, 0), (3, 1, 0)]),
    1: gamma_from_rows([(3, 0, -1), (2, 0, -1), (1, 0, 1), (0, 0, 1)]),
    2: gamma_from_rows([(3, -1, 0), (2, 1, 0), (1, 1, 0), (0, -1, 0)]),
    3: gamma_from_rows([(2, 0, -1), (3, 0, 1), (0, 0, 1), (1, 0, -1)]),
    4: gamma_from_rows([(2, -1, 0), (3, -1, 0), (0, -1, 0), (1, -1, 0)]),
    5: gamma_from_rows([(0, 1, 0), (1, 1, 0), (2, -1, 0), (3, -1, 0)]),
}
# direction mu (0=t,1=x,2=y,3=z) -> gamma index (iGammaOfMu, src/geometry/geometry_lx.hpp:240-244)
GMU = [GAMMA[4], GAMMA[1], GAMMA[2], GAMMA[3]]
G5 = GAMMA[5]


def cplx(a):
    """nissa array (..., 2) -> complex array (...)."""
    return a[..., 0] + 1j * a[..., 1]


def real(c):
    return np.ascontiguousarray(np.stack([c.real, c.imag], axis=-1))


def rel_l2(a, b):
    na = np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel())
    nb = np.linalg.norm(np.asarray(b).ravel())
    return na / nb if nb else na


def coords(L):
    T, X, Y, Z = L
    t, x, y, z = np.meshgrid(np.arange(T), np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    return np.stack([t.ravel(), x.ravel(), y.ravel(), z.ravel()], axis=1)  # lx order


def parity(L):
    return coords(L).sum(axis=1) % 2


def tm_conf(L, seed, eps=0.3, theta_t=1.0):
    """Synthetic gauge field with the antiperiodic time phase (put_boundaries_conditions)."""
    conf = nq.gauge_random_lx(L, seed, eps)
    nq.gauge_phase_lx(L, conf, [theta_t, 0, 0, 0])
    return conf


def st_conf(L, seed, eps=0.3):
    """Staggered links: eta phases and the antiperiodic time phase exp(-i pi/T)
    (add_or_rem_stagphases_to_conf, add_antiperiodic_condition_to_backfield)."""
    conf = nq.gauge_random_lx(L, seed, eps)
    nq.gauge_stagphases_lx(L, conf)
    nq.gauge_phase_lx(L, conf, [-1.0, 0, 0, 0])
    return conf


def random_su3_field(L, seed):
    """Site-local random SU(3) matrices (V,3,3,2), e.g. for gauge transformations."""
    V = int(np.prod(L))
    c = nq.gauge_random_lx(L, seed, 1.7)
    return np.ascontiguousarray(c[:, 0])


def gauss_eo(Vh, seed, ncomp=12):
    rng = np.random.default_rng(seed)
    shape = (Vh, 4, 3, 2) if ncomp == 12 else (Vh, 3, 2)
    return rng.normal(size=shape) * np.sqrt(0.5)
