"""Reference-named entries of the boundary beyond the data vector, on the GPU
against the oracle:
- the Limber integrands int_for_C_{ss,gs,gg,gk,ks}_tomo_limber and
  int_for_C_kk_limber (cosmo2D.h:116-130) at scale factors across each
  probe's range, through clh_int_for_limber and the single-point
  reference-named functions with the reference's parameter arrays: NLA with
  magnification, shifts and stretch; ss with deriv = 1 and the B mode; gg on
  the linear P(k,z); TATT; one-loop bias; the CMB lensing probes;
- the 6x2pt scalars w_gk_tomo, w_ks_tomo (cosmo2D.c:913, 1107), the C_gk /
  C_ks / C_kk tables (:3020, 3270, 3484) at and between their ln-ell nodes and
  the direct integrals (:2960, 3217, 3449).
Bars: integrands 1e-9 relative with a floor of 1e-6 of the largest value over
the probed a; C_ell and w(theta) 1e-6 (the FP64 parity metric)."""
import ctypes

import numpy as np
import pytest

from cosmolike_amd import Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu

ELLS = (2.0, 150.5, 2000.0)


def _close(got, ref, tol=1e-9):
    got, ref = np.asarray(got), np.asarray(ref)
    floor = 1e-6 * np.abs(ref).max()
    den = np.maximum(np.abs(ref), floor)
    den[den == 0] = 1.0
    return float((np.abs(got - ref) / den).max())


def _agrid(n=64):
    return np.linspace(0.2, 0.995, n)


def _nuis(p, seed=3):
    rng = np.random.default_rng(seed)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    if p.nl:
        p.lens_stretch = 1.0 + rng.normal(0, 0.03, p.nl)
        p.bmag = np.linspace(-0.8, 1.2, p.nl)
    return p


def _cases(p, probes):
    out = []
    for l in ELLS:
        if 0 in probes:
            out += [(0, [0, 3, l, 1, 0]), (0, [4, 4, l, 1, 1]), (0, [1, 2, l, 0, 0])]
        if 1 in probes:
            out += [(1, [0, 2, l, 0]), (1, [p.nl - 1, p.ns - 1, l, 0])]
        if 2 in probes:
            out += [(2, [1, 1, l, 0, 0]), (2, [p.nl - 1, p.nl - 1, l, 0, 0])]
    return out


def _check(c, o, cases, tol=1e-9):
    a = _agrid()
    worst = 0.0
    for probe, ar in cases:
        got = c.int_for(probe, ar, a)
        ref = o.int_for(probe, ar, a)
        e = _close(got, ref, tol)
        worst = max(worst, e)
        assert e < tol, (probe, ar, e)
    return worst


def test_integrands_nla():
    p = _nuis(synth.full_3x2pt_problem())
    c = Cosmolike().load(p)
    o = O.Oracle().load(p)
    print("worst", _check(c, o, _cases(p, (0, 1, 2))))
    # EE = 0 under NLA is zero
    assert np.all(c.int_for(0, [1, 2, 300.0, 0, 0], _agrid()) == 0.0)


def test_integrand_gg_linear_ps():
    p = _nuis(synth.nonlimber_problem())
    c = Cosmolike().load(p)
    o = O.Oracle().load(p)
    cases = [(2, [0, 0, l, 1, 0]) for l in ELLS] + [(2, [p.nl - 1, p.nl - 1, l, 1, 0]) for l in ELLS]
    print("worst", _check(c, o, cases))


def test_integrands_tatt():
    p = _nuis(synth.tatt_problem())
    c = Cosmolike().load(p)
    o = O.Oracle().load(p)
    print("worst", _check(c, o, _cases(p, (0, 1))))


def test_integrands_oneloop():
    p = _nuis(synth.oneloop_problem())
    c = Cosmolike().load(p)
    o = O.Oracle().load(p)
    cases = []
    for l in ELLS:
        cases += [(1, [1, 4, l, 1]), (1, [1, 4, l, 0]), (2, [2, 2, l, 0, 1]), (2, [2, 2, l, 0, 0])]
    print("worst", _check(c, o, cases))


@pytest.fixture(scope="module")
def six():
    cos = synth.random_draws(1, seed=7, ragged=False)[0][0]
    p = _nuis(synth.sixpt_problem(cos))
    return p, O.Oracle().load(p)


def test_integrands_cmb(six):
    p, o = six
    c = Cosmolike().load(p)
    cases = []
    for l in ELLS:
        cases += [(3, [0, l, 0]), (3, [p.nl - 1, l, 0]), (4, [0, l]), (4, [p.ns - 1, l]), (5, [l])]
    print("worst", _check(c, o, cases))


def test_reference_named_integrands(six):
    p, o = six
    d = Cosmolike.default()
    d.load(p)
    L = d.L
    for probe, name, ar in ((0, "int_for_C_ss_tomo_limber", [1, 3, 500.0, 1, 0]),
                            (1, "int_for_C_gs_tomo_limber", [2, 5, 500.0, 0]),
                            (2, "int_for_C_gg_tomo_limber", [2, 2, 500.0, 0, 0]),
                            (3, "int_for_C_gk_tomo_limber", [2, 500.0, 0]),
                            (4, "int_for_C_ks_tomo_limber", [5, 500.0]),
                            (5, "int_for_C_kk_limber", [500.0])):
        arr = (ctypes.c_double * len(ar))(*ar)
        fn = getattr(L, name)
        for a in (0.6, 0.85):
            ref = o.int_for(probe, ar, [a])[0]
            got = fn(a, arr)
            assert abs(got - ref) <= 1e-9 * max(abs(ref), 1e-300), (name, a, got, ref)
        assert np.isnan(fn(1.5, arr))  # a out of (0, 1)
    bad = (ctypes.c_double * 5)(0, 99, 500.0, 1, 0)
    assert np.isnan(L.int_for_C_ss_tomo_limber(0.5, bad))


def test_cmb_scalars(six):
    p, o = six
    d = Cosmolike.default()
    d.load(p)
    L = d.L
    gk, ks = o.w_gk(), o.w_ks()
    for b in range(p.nl):
        for t in (0, 9, p.ntheta - 1):
            assert abs(L.w_gk_tomo(t, b, 1) - gk[b, t]) <= 1e-6 * np.abs(gk[b]).max()
    for b in range(p.ns):
        for t in (0, 9, p.ntheta - 1):
            assert abs(L.w_ks_tomo(t, b, 1) - ks[b, t]) <= 1e-6 * np.abs(ks[b]).max()
    assert np.isnan(L.w_gk_tomo(0, p.nl, 1)) and np.isnan(L.w_ks_tomo(0, 0, 0))
    # tables: at their ln-ell nodes and halfway between, against the direct integrals
    lim0, lim1 = np.log(20.0), np.log(p.lmax + 1.0)
    dx = (lim1 - lim0) / (p.n_ell - 1)
    nodes = np.exp(lim0 + dx * np.array([3, 100, 400]))
    mids = np.exp(lim0 + dx * np.array([3.5, 100.5]))
    for probe, fn, nb in (("gk", L.C_gk_tomo_limber, p.nl), ("ks", L.C_ks_tomo_limber, p.ns),
                          ("kk", None, 1)):
        ref_n = o.cl_cmb_nointerp(probe, nodes)
        lo = o.cl_cmb_nointerp(probe, np.exp(lim0 + dx * np.array([3, 100])))
        hi = o.cl_cmb_nointerp(probe, np.exp(lim0 + dx * np.array([4, 101])))
        for b in range(nb):
            get = (lambda l: L.C_kk_limber(l)) if fn is None else (lambda l, b=b: fn(l, b))
            for i, l in enumerate(nodes):
                assert abs(get(l) - ref_n[b, i]) <= 1e-6 * abs(ref_n[b, i]), (probe, b, l)
            for i, l in enumerate(mids):
                ref = 0.5 * (lo[b, i] + hi[b, i])
                assert abs(get(l) - ref) <= 1e-6 * abs(ref), (probe, b, l)
    # direct integrals at arbitrary ell
    ells = [7.0, 333.3, 4100.0]
    for probe, fn, nb in (("gk", L.C_gk_tomo_limber_nointerp, p.nl), ("ks", L.C_ks_tomo_limber_nointerp, p.ns)):
        ref = o.cl_cmb_nointerp(probe, ells)
        for b in range(nb):
            for i, l in enumerate(ells):
                assert abs(fn(l, b, 0) - ref[b, i]) <= 1e-6 * abs(ref[b, i])
        assert fn(100.0, 0, 1) == 0.0
    ref = o.cl_cmb_nointerp("kk", ells)
    for i, l in enumerate(ells):
        assert abs(L.C_kk_limber_nointerp(l, 0) - ref[0, i]) <= 1e-6 * abs(ref[0, i])
    got = d.cl_nointerp("gk", ells)
    assert np.allclose(got, o.cl_cmb_nointerp("gk", ells), rtol=1e-6, atol=0)


def test_cmb_scalars_need_6x2pt():
    d = Cosmolike.default()
    d.load(_nuis(synth.full_3x2pt_problem()))
    assert np.isnan(d.L.C_kk_limber(100.0)) and np.isnan(d.L.w_gk_tomo(0, 0, 1))
