"""Pins of the oracle's 6x2pt CMB lensing probes (cosmolike/cosmo2D.c:
913-1296, 2857-3516; radial_weights.c:107-113; redshift_spline.c:1111-1125):

* C_kk, C_gk and C_ks against an independent continuum Limber integral with
  W_k = 3/2 Omega_m chi/a (chi* - chi)/chi*, chi* = chi(z = 1090);
* the CMB filter: zero filter range gives zero w_gk / w_ks, and w_gk is
  linear in the pixel window;
* the 6x2pt data vector: the 3x2pt blocks unchanged, w_ks times (1 + m_s),
  C_kk at the Fourier ells, and every block moved as a whole by ord.
"""
import math

import numpy as np
import pytest
from scipy import interpolate

from cosmolike_amd import synth

import oracle as O

COVERH0 = 2997.92458


@pytest.fixture(scope="module")
def six():
    p = synth.sixpt_problem()
    rng = np.random.default_rng(2)
    p.shear_m = rng.normal(0, 0.02, p.ns)
    return p, O.Oracle().load(p)


def _nz(table):
    z = table[:, 0]
    n = len(z)
    zmin, zmax = max(z[0], 1e-5), z[-1] + (z[-1] - z[0]) / (n - 1)
    dz = (zmax - zmin) / n
    zc = zmin + (np.arange(n) + 0.5) * dz
    out = []
    for b in range(table.shape[1] - 1):
        h = table[:, b + 1]
        cs = interpolate.CubicSpline(zc, h / (h.sum() * dz), bc_type="natural")
        out.append(lambda x, cs=cs, lo=zc[0], hi=zc[-1]: np.where((x >= lo) & (x <= hi), cs(x), 0.0))
    return out


def test_cmb_limber_vs_continuum(six):
    p, o = six
    lk, pz = p.log10k, p.pk_z
    rgi = interpolate.RegularGridInterpolator((lk, pz), p.lnP.reshape(len(lk), len(pz)), method="linear",
                                              bounds_error=False, fill_value=None)

    def P(k, z):
        return np.exp(rgi(np.c_[np.log10(k / COVERH0), z])) / COVERH0**3

    def chi(z):
        return np.interp(z, p.chi_z, p.chi) / COVERH0

    def E(z):
        return np.sqrt(p.omega_m * (1 + z) ** 3 + 1 - p.omega_m)

    chis = chi(1090.0)
    z = np.linspace(1e-4, 40.0, 200001)
    cz, Ez, a = chi(z), E(z), 1 / (1 + z)
    Wk = 1.5 * p.omega_m * cz / a * (chis - cz) / chis
    lens, src = _nz(p.lens), _nz(p.source)
    for l in (40.0, 400.0, 2500.0):
        L = l + 0.5
        pref = l * (l + 1) / L**2
        c = np.trapezoid(Wk**2 / cz**2 * P(L / cz, z) / Ez, z) * pref**2
        ref = o.cl_cmb_nointerp("kk", [l])[0, 0]
        assert abs(ref / c - 1) < 2e-3, ("kk", l, ref, c)
        for zl in (0, 3):
            w = lens[zl](z) * Ez * p.b1[zl]  # bmag = 0
            c = np.trapezoid(Wk * w / cz**2 * P(L / cz, z) / Ez, z) * pref
            ref = o.cl_cmb_nointerp("gk", [l])[zl, 0]
            assert abs(ref / c - 1) < 2e-3, ("gk", zl, l, ref, c)
    # ks: W_kappa of the source bin (its lensing efficiency) times W_k, IA off
    q = synth.sixpt_problem()
    q.ia_A = 0.0
    oq = O.Oracle().load(q)
    zg = np.linspace(1e-4, 3.0, 30001)
    czg, Ezg, ag = chi(zg), E(zg), 1 / (1 + zg)
    for zs in (2, 8):
        nz = src[zs](zg)
        inv = np.where(czg > 0, nz / np.maximum(czg, 1e-30), 0.0)
        A = np.concatenate([[0], np.cumsum(0.5 * (nz[1:] + nz[:-1]) * np.diff(zg))])
        B = np.concatenate([[0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(zg))])
        g = (A[-1] - A) - czg * (B[-1] - B)
        Wks = 1.5 * p.omega_m * czg / ag * g
        Wkg = 1.5 * p.omega_m * czg / ag * (chis - czg) / chis
        for l in (40.0, 400.0):
            L = l + 0.5
            pref = l * (l + 1) / L**2 * math.sqrt((l - 1) * l * (l + 1) * (l + 2)) / L**2
            c = np.trapezoid(Wks * Wkg / czg**2 * P(L / czg, zg) / Ezg, zg) * pref
            ref = oq.cl_cmb_nointerp("ks", [l])[zs, 0]
            assert abs(ref / c - 1) < 3e-3, ("ks", zs, l, ref, c)


def test_cmb_filter(six):
    p, o = six
    base = o.w_gk()
    o.set_cmb(p.cmb_fwhm, p.cmb_lmin, p.cmb_lmax, 2.0 * p.cmb_win)
    assert np.allclose(o.w_gk(), 2.0 * base, rtol=1e-12, atol=0)
    o.set_cmb(p.cmb_fwhm, 100, 50, p.cmb_win)  # empty filter range
    assert np.all(o.w_gk() == 0) and np.all(o.w_ks() == 0)
    o.set_cmb(p.cmb_fwhm, p.cmb_lmin, p.cmb_lmax, p.cmb_win)


def test_6x2pt_assembly(six):
    p, o = six
    ells = synth.ells_fourier(p.kk_nells, p.kk_lmin, p.kk_lmax)
    dv6 = o.datavector_6x2pt(ells)
    dv3 = o.datavector()
    n3, nt = len(dv3), p.ntheta
    assert np.array_equal(dv6[:n3], dv3)
    gk, ks, kk = o.w_gk(), o.w_ks(), o.cl_kk(ells)
    assert np.array_equal(dv6[n3:n3 + p.nl * nt], gk.reshape(-1))
    got_ks = dv6[n3 + p.nl * nt:n3 + (p.nl + p.ns) * nt].reshape(p.ns, nt)
    assert np.allclose(got_ks, ks * (1 + p.shear_m)[:, None], rtol=1e-15, atol=0)
    assert np.array_equal(dv6[-len(ells):], kk)
    # reversed block order: every block moves whole
    rev = o.datavector_6x2pt(ells, ord=(5, 4, 3, 2, 1, 0))
    sizes = [2 * 55 * nt, 50 * nt, 5 * nt, 5 * nt, 10 * nt, len(ells)]
    starts = np.concatenate([[0], np.cumsum(sizes)])
    pos = 0
    for b in range(5, -1, -1):
        assert np.array_equal(rev[pos:pos + sizes[b]], dv6[starts[b]:starts[b] + sizes[b]]), b
        pos += sizes[b]


def test_kk_bandpower(six):
    """init_cmb_auto_bandpower: the kk block becomes sum_L M[j][L - lmin]
    C_kk(L) - offset[j] (generic_interface.hpp:1095-1118). One-hot rows give
    C_kk at single multipoles on both sides of LMIN_tab (direct at L <= 20,
    the table above); the map is linear in M; the other blocks stay as they
    are and the block size is the number of band powers."""
    p, o = six
    ells = synth.ells_fourier(p.kk_nells, p.kk_lmin, p.kk_lmax)
    dv6 = o.datavector_6x2pt(ells)
    n5 = len(dv6) - len(ells)
    lmin, lmax = 8, 400
    picks = [8, 15, 20, 21, 22, 57, 400]
    M = np.zeros((len(picks), lmax - lmin + 1))
    for j, L in enumerate(picks):
        M[j, L - lmin] = 1.0
    off = np.linspace(-1e-9, 3e-9, len(picks))
    o.set_kk_bandpower(len(picks), lmin, lmax, M.reshape(-1), off)
    try:
        got = o.datavector_6x2pt(ells)
        assert len(got) == n5 + len(picks)
        assert np.array_equal(got[:n5], dv6[:n5])
        ref = np.array([o.cl_kk([float(L)])[0] for L in picks]) - off
        assert np.allclose(got[n5:], ref, rtol=1e-14, atol=0)
        # direct below LMIN_tab, the table above: these agree only closely
        direct = o.cl_cmb_nointerp("kk", [21.0, 57.0])[0]
        assert np.allclose(got[n5 + 3:n5 + 6:2] + off[3:6:2], direct, rtol=2e-3)
        # linear in M: a normalised band equals the weighted one-hot sum
        nbp, lo, hi, Mb, offb = synth.kk_bandpower_binning()
        o.set_kk_bandpower(nbp, lo, hi, Mb, offb)
        bp = o.datavector_6x2pt(ells)[n5:]
        Ls = np.arange(lo, hi + 1.0)
        ckk = np.array([o.cl_kk([L])[0] for L in Ls])
        assert np.allclose(bp, Mb.reshape(nbp, -1) @ ckk - offb, rtol=1e-12, atol=0)
        assert np.all(bp > 0)
    finally:
        o.set_kk_bandpower(0, 0, 0, np.zeros(1))
    assert np.array_equal(o.datavector_6x2pt(ells), dv6)
