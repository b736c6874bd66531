"""HIP path vs the FP64 CPU oracle on the same seeded inputs.

Bar: the plain FP64 path matches the oracle to 1e-6 relative.  The relative
error of an entry is |gpu - ref| / max(|ref|, 1e-3 * max|ref| of its
(probe, pair) row), so zero crossings of xi-, gamma_t and w do not divide by
~0; everywhere else it is the plain relative error.
"""
import numpy as np
import pytest

from cosmolike_amd import Cosmolike, LEGENDRE_FOLDED, LEGENDRE_FULL, synth

import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def rel_err(gpu, ref):
    gpu = np.atleast_2d(gpu)
    ref = np.atleast_2d(ref)
    floor = 1e-3 * np.abs(ref).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(ref), floor)
    den[den == 0] = 1.0
    return float((np.abs(gpu - ref) / den).max())


@pytest.fixture(scope="module")
def shear():
    p = synth.shear_only_problem()
    o = O.Oracle().load(p)
    xp, xm = o.xi_pm()
    ss, _, _ = o.cl_tables()
    return p, o, xp, xm, ss


@pytest.fixture(scope="module")
def full():
    p = synth.full_3x2pt_problem()
    rng = np.random.default_rng(7)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.lens_stretch = 1.0 + rng.normal(0, 0.02, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    o = O.Oracle().load(p)
    dv = o.datavector()
    tabs = o.cl_tables()
    return p, o, dv, tabs


def test_shear_cl_tables(shear):
    p, o, xp, xm, ss = shear
    c = Cosmolike().load(p)
    g_ss, _, _ = c.cl_tables(p.n_ell)
    assert rel_err(g_ss, ss) < RTOL


def test_shear_xi_folded(shear):
    p, o, xp, xm, ss = shear
    c = Cosmolike().load(p, LEGENDRE_FOLDED)
    gp, gm = c.xi_pm()
    print("xi+ err", rel_err(gp, xp), "xi- err", rel_err(gm, xm))
    assert rel_err(gp, xp) < RTOL
    assert rel_err(gm, xm) < RTOL


def test_shear_xi_full(shear):
    p, o, xp, xm, ss = shear
    c = Cosmolike().load(p, LEGENDRE_FULL)
    gp, gm = c.xi_pm()
    print("xi+ err", rel_err(gp, xp), "xi- err", rel_err(gm, xm))
    assert rel_err(gp, xp) < RTOL
    assert rel_err(gm, xm) < RTOL


def test_shear_fourier_nointerp(shear):
    # config-1 Fourier output: C_ell at 20 log-spaced ell in [20, 3000]
    p, o, *_ = shear
    ells = np.exp(np.linspace(np.log(20.0), np.log(3000.0), 20))
    ref = o.cl_nointerp("ss", ells)
    c = Cosmolike().load(p)
    assert rel_err(c.cl_nointerp("ss", ells), ref) < RTOL


def _split(dv, p):
    nt, nss, ngs, nl = p.ntheta, p.ns * (p.ns + 1) // 2, p.ns * p.nl, p.nl
    a = nt * nss
    return (dv[:a].reshape(nss, nt), dv[a:2 * a].reshape(nss, nt),
            dv[2 * a:2 * a + nt * ngs].reshape(ngs, nt), dv[2 * a + nt * ngs:].reshape(nl, nt))


@pytest.mark.parametrize("mode", [LEGENDRE_FOLDED, LEGENDRE_FULL])
def test_3x2pt_datavector(full, mode):
    p, o, dv, tabs = full
    c = Cosmolike().load(p, mode)
    g = c.compute_3x2pt_real()
    assert g.shape == dv.shape
    for name, a, b in zip(["xi+", "xi-", "gammat", "w"], _split(g, p), _split(dv, p)):
        e = rel_err(a, b)
        print(name, e)
        assert e < RTOL, name


def test_3x2pt_tables(full):
    p, o, dv, (ss, gs, gg) = full
    c = Cosmolike().load(p)
    a, b, d = c.cl_tables(p.n_ell)
    assert rel_err(a, ss) < RTOL
    assert rel_err(b, gs) < RTOL
    assert rel_err(d, gg) < RTOL


def test_3x2pt_order(full):
    p, o, dv, tabs = full
    c = Cosmolike().load(p)
    for ord in [(1, 0, 2), (2, 1, 0), (0, 2, 1)]:
        ref = o.datavector(ord)
        assert rel_err(c.compute_3x2pt_real(ord), ref) < RTOL


def test_low_ell_direct(full):
    p, o, *_ = full
    c = Cosmolike().load(p)
    ells = np.arange(1.0, 20.0)
    for probe in ("ss", "gs", "gg"):
        assert rel_err(c.cl_nointerp(probe, ells), o.cl_nointerp(probe, ells)) < RTOL, probe


def test_batch_matches_single(full):
    p, o, dv, tabs = full
    import torch
    draws = []
    for i, (cos, *_rest) in enumerate(synth.random_draws(3, seed=11, ragged=False)):
        q = synth.full_3x2pt_problem(cos)
        q.shear_m = np.full(q.ns, 0.01 * i)
        draws.append(q)
    c = Cosmolike().load(p)
    c.batch_upload(draws)
    n = c.datavector_size()
    out = torch.empty(len(draws) * n, dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    got = out.cpu().numpy().reshape(len(draws), n)
    for i, q in enumerate(draws):
        ref = O.Oracle().load(q).datavector()
        assert rel_err(got[i], ref) < RTOL, i


def test_reference_scalar_entries(full):
    p, o, dv, tabs = full
    c = Cosmolike.default()
    c.load(p)
    L = c.L
    xp, xm = o.xi_pm()
    gt = o.gammat()
    w = o.wtheta()
    nss = p.ns * (p.ns + 1) // 2
    # N_shear(ni, nj) is symmetric
    q = 0
    for i in range(p.ns):
        for j in range(i, p.ns):
            for nt in (0, 13, 25):
                assert abs(L.xi_pm_tomo(1, nt, j, i, 1) / xp[q, nt] - 1) < RTOL
                assert abs(L.xi_pm_tomo(-1, nt, i, j, 1) / xm[q, nt] - 1) < 10 * RTOL or \
                    abs(L.xi_pm_tomo(-1, nt, i, j, 1) - xm[q, nt]) < RTOL * np.abs(xm[q]).max()
            q += 1
    assert q == nss
    for zl in range(p.nl):
        for zs in range(p.ns):
            v = L.w_gammat_tomo(5, zl, zs, 1)
            assert abs(v - gt[zl * p.ns + zs, 5]) <= RTOL * np.abs(gt[zl * p.ns + zs]).max()
        assert abs(L.w_gg_tomo(7, zl, zl, 1) - w[zl, 7]) <= RTOL * np.abs(w[zl]).max()
    # C_*_tomo_limber: table interpolation at arbitrary ell
    ss, gs, gg = tabs
    lim0 = np.log(19.0)
    lims = np.linspace(lim0, np.log(50001.0), p.n_ell)
    for l in (25.0, 300.5, 4000.0):
        ref = np.interp(np.log(l), lims, ss[3])
        assert abs(L.C_ss_tomo_limber(l, 0, 3, 1) / ref - 1) < RTOL
    assert np.isnan(L.xi_pm_tomo(1, 0, 0, 99, 1))
    assert np.isnan(L.w_gg_tomo(0, 0, 1, 1))
    assert np.isnan(L.w_gg_tomo(0, 0, 0, 2))  # limber is 0 or 1


def test_error_paths_leave_output(full):
    p, *_ = full
    c = Cosmolike().load(p)
    n = c.datavector_size()
    import ctypes
    out = np.full(n, 123.0)
    o = (ctypes.c_int * 3)(0, 1, 2)
    rc = c.L.clh_compute_3x2pt_real(c.h, o, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n - 1)
    assert rc == -1 and np.all(out == 123.0)
    bad = p.chi.copy()
    bad[5] = np.nan
    with pytest.raises(Exception):
        c.set_distances(p.chi_z, bad)
    # a chi(z) table too short for chi(a_min) at z=40 is a domain error, not an over-read
    c.set_distances(p.chi_z[:420], p.chi[:420])
    rc = c.L.clh_compute_3x2pt_real(c.h, o, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    assert rc == -4 and np.all(out == 123.0)


def test_large_tables_not_staged():
    # chi(z) and G(z) tables too large for the gg kernels' LDS staging (the
    # kernels then walk the tables in global memory): same bar
    p = synth.full_3x2pt_problem()
    z = np.linspace(0.0, p.chi_z[-1], 3000)
    p.chi_z, p.chi = z, np.interp(z, p.chi_z, p.chi)
    p.G_z, p.G = z, np.interp(z, p.G_z, p.G)
    ref = O.Oracle().load(p).datavector()
    got = Cosmolike().load(p).compute_3x2pt_real()
    print("err", rel_err(got, ref))
    assert rel_err(got, ref) < RTOL
