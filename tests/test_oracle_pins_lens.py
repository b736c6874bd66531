"""Pins of the oracle's lens-sample side: the photo-z shift and stretch of
nz_lens_photoz, zmean, the magnification lens efficiency g_lens with its
pre-integral, and the gs / gg Limber integrals with magnification bias,
source and lens shifts and lens stretch.

Each reference value is built independently of the oracle's code: the n(z)
spline from scipy's natural CubicSpline of the normalised histogram, the
shift / stretch as the transform of its argument, integrals as dense
trapezoids in z of the continuum formulas.  The lens bins are smooth bumps
that vanish with their first derivative at compact edges, so that the
integration limits matter: a limit that ignores the stretch or the 2|dz|
margin (redshift_spline.c:50-79), or a g_lens without the pre-integral from
the source sample's a_min (:1052-1071), cuts off a percent of the integral.
"""
import math

import numpy as np
import pytest
from scipy import interpolate

from cosmolike_amd import synth

import oracle as O

COVERH0 = 2997.92458
LENS_ZMAX = 1.5  # lens histogram on [0, 1.5): shorter than the source sample's [0, 3)


def bump_lens_table(bins, nrows=150, zmax=LENS_ZMAX):
    """(z_left, n_1, ...) rows; n_b = (z-lo)^2 (hi-z)^2 on (lo, hi), else 0."""
    zl = np.linspace(0.0, zmax, nrows + 1)[:-1]
    zc = zl + 0.5 * zmax / nrows
    out = np.zeros((nrows, len(bins) + 1))
    out[:, 0] = zl
    for i, (lo, hi) in enumerate(bins):
        m = (zc > lo) & (zc < hi)
        out[m, i + 1] = (zc[m] - lo) ** 2 * (hi - zc[m]) ** 2
    return out


BINS = [(0.15, 0.55), (0.4, 0.85), (1.12, 1.46)]


def lens_problem(bmag, lens_dz, stretch, src_dz=None, bins=BINS):
    p = synth.full_3x2pt_problem()
    p.lens = bump_lens_table(bins)
    nl = len(bins)
    p.b1 = np.array([1.3, 1.7, 2.1])
    p.bmag = np.asarray(bmag, dtype=float)
    p.lens_dz = np.asarray(lens_dz, dtype=float)
    p.lens_stretch = np.asarray(stretch, dtype=float)
    p.pm = np.zeros(nl)
    if src_dz is not None:
        p.src_dz = np.asarray(src_dz, dtype=float)
    return p


def _spline(table, b):
    """The normalised n(z) spline of bin b on the histogram centres (zero
    outside the first and last centre) and the support [zmin, zmax]."""
    z = table[:, 0]
    n = len(z)
    zmin_all = max(z[0], 1e-5)
    zmax_all = z[-1] + (z[-1] - z[0]) / (n - 1)
    dz = (zmax_all - zmin_all) / n
    zc = zmin_all + (np.arange(n) + 0.5) * dz
    h = table[:, b + 1]
    y = np.array([h[int(np.floor((x - z[0]) / ((z[-1] - z[0]) / (n - 1))))] for x in zc])
    y = y / (y.sum() * dz)
    cs = interpolate.CubicSpline(zc, y, bc_type="natural")
    nz = np.nonzero(h > 0.999e-8 * h.max())[0]

    def f(x):
        x = np.asarray(x, dtype=float)
        return np.where((x > zc[0]) & (x < zc[-1]), cs(x), 0.0)
    return f, z[nz[0]], z[nz[-1]], zmax_all


def _lens_fn(table, b, dz, st):
    f, lo, hi, _ = _spline(table, b)
    zz = np.linspace(lo, hi, 200001)
    fz = f(zz)
    zbar = np.trapezoid(zz * fz, zz) / np.trapezoid(fz, zz)
    return (lambda z: f((np.asarray(z) - dz - zbar) / st + zbar) / st), zbar


def _background(p):
    G0 = np.interp(0.0, p.G_z, p.G)
    chi = lambda z: np.interp(z, p.chi_z, p.chi) / COVERH0
    E = lambda z: np.sqrt(p.omega_m * (1 + z) ** 3 + 1 - p.omega_m)
    D = lambda z: np.interp(z, p.G_z, p.G) / (1 + z) / G0
    lk, pz = p.log10k, p.pk_z
    rgi = interpolate.RegularGridInterpolator((lk, pz), p.lnP.reshape(len(lk), len(pz)), method="linear",
                                              bounds_error=False, fill_value=None)
    P = lambda k_c, z: np.exp(rgi(np.c_[np.log10(k_c / COVERH0), z])) / COVERH0**3
    return chi, E, D, P


def _efficiency(n, chi, z, ztop):
    """g(z) = int_z^ztop n(z') (1 - chi(z)/chi(z')) dz' on the grid z (z[-1] = ztop)."""
    nz = n(z)
    cz = chi(z)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(cz > 0, nz / cz, 0.0)
    A = np.concatenate([[0], np.cumsum(0.5 * (nz[1:] + nz[:-1]) * np.diff(z))])
    B = np.concatenate([[0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(z))])
    return (A[-1] - A) - cz * (B[-1] - B)


def test_lens_shift_stretch_and_zmean():
    """nz_lens_photoz (redshift_spline.c:736-762): n((z - dz - zbar)/s + zbar)/s,
    zero outside the knots; zmean (:794-844) at zero shift and unit stretch."""
    p = lens_problem([0.0] * 3, [0.04, -0.03, 0.05], [1.3, 0.8, 1.1])
    o = O.Oracle().load(p)
    zm = o.zmean_lens()
    z = np.linspace(0.0, 2.0, 4001)
    for b in range(3):
        fn, zbar = _lens_fn(p.lens, b, p.lens_dz[b], p.lens_stretch[b])
        assert abs(zm[b] - zbar) < 1e-7, (b, zm[b], zbar)
        ref = fn(z)
        got = o.nz_lens(b, z)
        assert np.allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max()), b
    # unshifted, unstretched: zmean is the same (set_lens_sample evaluates it so)
    q = lens_problem([0.0] * 3, [0.0] * 3, [1.0] * 3)
    assert np.allclose(O.Oracle().load(q).zmean_lens(), zm, rtol=0, atol=1e-14)


def test_g_lens_vs_direct_integral():
    """g_lens at its coarse-table nodes = int_{z(a)}^{zmax,src} n_lens(z')
    (1 - chi(a)/chi(z')) dz' (redshift_spline.c:1014-1109): the table spans the
    lens sample's a range, and the part of n_lens between the lens and the
    source sample's zmax enters through the pre-integral."""
    # the third bin reaches the top of the lens sample and is shifted up
    p = lens_problem([0.5, 0.5, 0.5], [0.02, -0.03, 0.1], [1.2, 0.9, 1.0],
                     bins=[(0.15, 0.55), (0.4, 0.85), (1.2, 1.5)])
    o = O.Oracle().load(p)
    chi = lambda z: np.interp(z, p.chi_z, p.chi)
    _, _, _, zmax_lens = _spline(p.lens, 0)
    ztop = 3.0  # the source sample's zmax
    amin = 1.0 / (zmax_lens + 1.0)
    dac = (0.999999 - amin) / 224
    # the third bin, shifted up, reaches above the lens sample's zmax
    fn2, _ = _lens_fn(p.lens, 2, p.lens_dz[2], p.lens_stretch[2])
    zz = np.linspace(zmax_lens, ztop, 20001)
    tail = np.trapezoid(fn2(zz), zz)
    assert tail > 0.1
    for b in range(3):
        fn, _ = _lens_fn(p.lens, b, p.lens_dz[b], p.lens_stretch[b])
        for k in (1, 30, 80, 140, 200):
            a = amin + k * dac
            zk = 1 / a - 1
            zg = np.linspace(zk, ztop, 60001)
            ref = np.trapezoid(fn(zg) * (1 - chi(zk) / chi(zg)), zg)
            got = o.g_lens(b, [a])[0]
            scale = max(abs(ref), 1e-3 * np.trapezoid(fn(zg), zg), 1e-12)
            assert abs(got - ref) < 2e-4 * scale, (b, k, got, ref)
    # without the pre-integral the high bin's efficiency at the table's first
    # node would miss the tail above the lens sample's zmax
    a0 = amin + dac
    z0 = 1 / a0 - 1
    zg = np.linspace(z0, zmax_lens, 20001)
    short = np.trapezoid(fn2(zg) * (1 - chi(z0) / chi(zg)), zg)
    got = o.g_lens(2, [a0])[0]
    assert abs(got - short) > 0.02 * abs(got)


@pytest.mark.parametrize("case", ["shifted", "magnified"])
def test_gs_gg_vs_continuum_with_lens_nuisance(case):
    """C_gs and C_gg (cosmo2D.c:1890-2073, 2490-2666) against continuum Limber
    integrals over the whole z range, with source and lens shifts, lens
    stretch and (case "magnified") magnification bias: the integration limits
    amin_lens / amax_lens must cover the shifted, stretched bins and, with
    b_mag != 0, reach down to the source sample's zmin."""
    rng = np.random.default_rng(9)
    # bin 0: stretched, not shifted (its limits then depend on the stretch alone)
    if case == "shifted":
        p = lens_problem([0.0] * 3, [0.0, -0.04, 0.03], [1.3, 0.8, 1.15], src_dz=rng.normal(0, 0.02, 10))
    else:
        p = lens_problem([1.6, -0.9, 2.2], [0.0, 0.05, -0.05], [1.25, 1.1, 0.85], src_dz=rng.normal(0, 0.02, 10))
    o = O.Oracle().load(p)
    chi, E, D, P = _background(p)
    z = np.linspace(1e-4, 3.0, 60001)
    cz, Ez = chi(z), E(z)
    a = 1 / (1 + z)
    src = [_spline(p.source, b)[0] for b in range(p.ns)]
    srcs = [lambda zz, f=f, d=d: f(np.asarray(zz) - d) for f, d in zip(src, p.src_dz)]
    C1 = 0.5 * p.omega_m * 0.01389 / D(z)

    def q_shear(b):
        g = _efficiency(srcs[b], chi, z, 3.0)
        return 1.5 * p.omega_m * cz / a * g - C1 * srcs[b](z) * Ez

    lens = [_lens_fn(p.lens, b, p.lens_dz[b], p.lens_stretch[b])[0] for b in range(3)]
    ells = np.array([100.0, 1000.0, 5000.0])
    ref_gs = o.cl_nointerp("gs", ells)
    ref_gg = o.cl_nointerp("gg", ells)
    worst = 0.0
    for zl in range(3):
        wmag = 1.5 * p.omega_m * cz / a * _efficiency(lens[zl], chi, z, 3.0)
        ngal = lens[zl](z) * Ez
        for n, l in enumerate(ells):
            L = l + 0.5
            pref = l * (l + 1) / L**2
            gal = ngal * p.b1[zl] + wmag * pref * p.bmag[zl]
            Pz = P(L / cz, z)
            c_gg = np.trapezoid(gal**2 / cz**2 * Pz / Ez, z)
            e = abs(ref_gg[zl, n] / c_gg - 1)
            worst = max(worst, e)
            assert e < 2e-3, ("gg", case, zl, l, ref_gg[zl, n], c_gg)
            for zs in (2, 9):
                # relative to the integral of |integrand| where the lensing and
                # IA terms of an overlapping pair cancel
                f = q_shear(zs) * gal / cz**2 * Pz / Ez * math.sqrt((l - 1) * l * (l + 1) * (l + 2)) / L**2
                c_gs = np.trapezoid(f, z)
                e = abs(ref_gs[zl * p.ns + zs, n] - c_gs) / max(abs(c_gs), np.trapezoid(np.abs(f), z))
                worst = max(worst, e)
                assert e < 2e-3, ("gs", case, zl, zs, l, ref_gs[zl * p.ns + zs, n], c_gs)
    print(case, "worst relative difference", worst)
