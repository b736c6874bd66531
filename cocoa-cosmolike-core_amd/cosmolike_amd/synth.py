"""Synthetic inputs for the tomographic two-point projection path.

Everything here is plain numpy FP64 and seeded, so the oracle and the HIP path
see bit-identical inputs.  The shapes follow what Cocoa hands to CosmoLike's
setters (cosmolike/generic_interface.cpp):

* ``set_distances(z, chi)``                       chi in Mpc/h      (:1085)
* ``set_growth(z, G)``                            G = D/a           (:1149)
* ``set_non_linear_power_spectrum(log10k, z, lnP)`` k-major lnP    (:1276, :1343)
* ``set_source_sample`` / ``set_lens_sample``     histogram tables whose
  column 0 is z and columns 1..N are n_i(z)      (:1907, :1802)

Cosmology: flat LCDM, P(k,z) = A k^ns T_BBKS(k)^2 D(z)^2 with A fixed by sigma8.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
from scipy import integrate, special

C_OVER_H0 = 2997.92458  # Mpc/h, cosmolike/structs.c:119


@dataclasses.dataclass
class Cosmology:
    omega_m: float = 0.3
    h: float = 0.7
    ns: float = 0.965
    sigma8: float = 0.8


def _E(z, om):
    return np.sqrt(om * (1.0 + z) ** 3 + (1.0 - om))


def background_grid(cmb: bool = False) -> np.ndarray:
    """z grid for chi(z) and G(z): dense below z=4, sparser up to z=50.

    chi_all reads chi[j+2] (cosmolike/cosmo3D.c:66-67) and the gg RSD term
    queries chi(a_min) at z=40 (cosmo2D.c:2540, structs.c:41), so the grid has
    to reach well past z=40.
    """
    z1 = np.linspace(0.0, 4.0, 401)
    z2 = np.geomspace(4.05, 50.0, 120)
    if not cmb:
        return np.concatenate([z1, z2])
    # CMB lensing (g_cmb, redshift_spline.c:1111-1125) reads chi(z = 1090)
    z3 = np.geomspace(52.0, 1300.0, 80)
    return np.concatenate([z1, z2, z3])


def comoving_distance(z: np.ndarray, om: float) -> np.ndarray:
    """chi(z) in Mpc/h for flat LCDM (high-accuracy cumulative Simpson)."""
    zf = np.linspace(0.0, float(z.max()), 200001)
    inv = 1.0 / _E(zf, om)
    cum = integrate.cumulative_simpson(inv, x=zf, initial=0.0)
    return C_OVER_H0 * np.interp(z, zf, cum)


def growth_D(z: np.ndarray, om: float) -> np.ndarray:
    """Linear growth D(z) normalised to D(0)=1: D(a) ~ E(a) int_0^a da'/(a' E(a'))^3
    (Heath 1977), by cumulative Simpson on a fine grid in a."""
    af = np.linspace(0.0, 1.0, 400001)
    e = np.empty_like(af)
    e[1:] = np.sqrt(om / af[1:] ** 3 + (1.0 - om))
    e[0] = np.inf
    integrand = np.zeros_like(af)
    integrand[1:] = 1.0 / (af[1:] * e[1:]) ** 3
    cum = integrate.cumulative_simpson(integrand, x=af, initial=0.0)
    d = np.zeros_like(af)
    d[1:] = 2.5 * om * e[1:] * cum[1:]
    a = 1.0 / (1.0 + np.asarray(z, dtype=np.float64))
    da = np.interp(a, af, d)
    return da / d[-1]


def t_bbks(k_hmpc: np.ndarray, om: float, h: float) -> np.ndarray:
    q = k_hmpc / (om * h)
    x = 2.34 * q
    return (np.log1p(x) / x) * (1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3
                                + (6.71 * q) ** 4) ** -0.25


def sigma8_unnorm(om: float, h: float, ns: float) -> float:
    lk = np.linspace(-5.0, 3.0, 40001)
    k = 10.0**lk
    x = 8.0 * k
    w = 3.0 * (np.sin(x) - x * np.cos(x)) / x**3
    pk = k**ns * t_bbks(k, om, h) ** 2
    integrand = k**3 * pk * w * w / (2.0 * np.pi**2) * np.log(10.0)
    return math.sqrt(integrate.simpson(integrand, x=lk))


def fpt_ia_tables(cosmo: Cosmology, n: int = 1100, kmin: float = 1e-4, kmax: float = 1e2):
    """Synthetic stand-ins for the FAST-PT IA power spectra, in set_IA_PS form:
    PS[12][n] on k_i = exp(ln kmin + i (ln kmax - ln kmin) / n) (h/Mpc), rows
    0..9 the tt E/B, ta dE1/dE2/0E0E/0B0B and mix A/B/D_EE/D_BB terms in
    (Mpc/h)^3 (smooth functions of the linear P(k) at z = 0 with distinct
    amplitudes and signs), row 10 = k, row 11 = P_lin.  They exercise every
    term of the TATT integrands; they are not one-loop integrals."""
    lk = np.log(kmin) + np.arange(n) * (np.log(kmax) - np.log(kmin)) / n
    k = np.exp(lk)
    amp = (cosmo.sigma8 / sigma8_unnorm(cosmo.omega_m, cosmo.h, cosmo.ns)) ** 2
    plin = amp * k**cosmo.ns * t_bbks(k, cosmo.omega_m, cosmo.h) ** 2
    shape = plin * (k / (k + 0.3)) ** 2
    coef = [0.8, 0.3, -0.5, 0.25, 0.6, 0.2, 0.4, -0.35, 0.5, 0.15]
    out = np.empty((12, n))
    for t, c in enumerate(coef):
        out[t] = c * shape * (1.0 + 0.1 * t * np.tanh(np.log(k / 0.1)))
    out[10] = k
    out[11] = plin
    return np.ascontiguousarray(out.reshape(-1)), kmin, kmax


def fpt_bias_tables(cosmo: Cosmology, n: int = 1100, kmin: float = 1e-4, kmax: float = 1e2):
    """Synthetic stand-ins for the FAST-PT one-loop bias spectra, in
    set_bias_PS form: PS[12][n] on k_i = exp(ln kmin + i (ln kmax - ln kmin) / n)
    (h/Mpc), rows 0..5 Pd1d2, Pd2d2, Pd1s2, Pd2s2, Ps2s2, Pd1p3 in (Mpc/h)^3
    (smooth functions of P_lin(k, z = 0) of tens of per cent of it, distinct
    amplitudes and signs),
    row 6 = k, row 7 = P_lin, rows 8..11 zero; and sigma4 = Pd2d2(kmin) / 2 as
    get_FPT_bias sets it.  They exercise every term of the one-loop
    integrands; they are not one-loop integrals."""
    lk = np.log(kmin) + np.arange(n) * (np.log(kmax) - np.log(kmin)) / n
    k = np.exp(lk)
    amp = (cosmo.sigma8 / sigma8_unnorm(cosmo.omega_m, cosmo.h, cosmo.ns)) ** 2
    plin = amp * k**cosmo.ns * t_bbks(k, cosmo.omega_m, cosmo.h) ** 2
    shape = plin * (k / (k + 0.2)) ** 2
    coef = [0.5, 0.3, -0.35, 0.2, 0.15, -0.12]
    out = np.zeros((12, n))
    for t, c in enumerate(coef):
        out[t] = c * shape * (1.0 + 0.15 * t * np.tanh(np.log(k / 0.2)))
    out[1] += 2.0 * 40.0  # Pd2d2 tends to 2 sigma4 at low k
    out[6] = k
    out[7] = plin
    return np.ascontiguousarray(out.reshape(-1)), kmin, kmax, 0.5 * out[1][0]


def oneloop_problem(cosmo: Cosmology | None = None, n_ell: int = 500, b3=None, bk=None) -> Problem:
    """The full 3x2pt layout with one-loop galaxy bias: b2 from b1
    (Lazeyras et al. form, as B2_FROM_B1 in cosmolike/bias.c:85-91),
    bs2 = -4/7 (b1 - 1), b3 and bK per lens bin, and synthetic FAST-PT bias
    tables."""
    cosmo = cosmo or Cosmology()
    p = full_3x2pt_problem(cosmo, n_ell)
    b1 = p.b1
    p.b2 = 0.412 - 2.143 * b1 + 0.929 * b1**2 + 0.008 * b1**3
    p.b3 = np.asarray(b3, dtype=np.float64) if b3 is not None else 0.3 * (b1 - 1.0)
    # bK multiplies k^2 P with k in H0/c units (k = 3000 at 1 h/Mpc)
    p.bk = np.asarray(bk, dtype=np.float64) if bk is not None else np.linspace(-4e-8, 4e-8, len(b1))
    p.bias_ps, p.bias_ps_kmin, p.bias_ps_kmax, p.bias_ps_sigma4 = fpt_bias_tables(cosmo)
    return p


def ells_fourier(nells: int, lmin: float, lmax: float) -> np.ndarray:
    """like.ell of init_binning_fourier (generic_interface.cpp:425-470): the
    geometric centres of nells log-spaced bins in [lmin, lmax]."""
    e = lmin * (lmax / lmin) ** (np.arange(nells + 1) / nells)
    return np.sqrt(e[:-1] * e[1:])


def healpix_window_like(nside: int = 2048, lmax: int = 6000) -> np.ndarray:
    """A smooth stand-in for a HEALPix pixel window (ell = 0..lmax): a
    Gaussian of the pixel scale, not the HEALPix tabulation."""
    l = np.arange(lmax + 1.0)
    return np.exp(-0.5 * (l / (2.0 * nside)) ** 2)


def sixpt_problem(cosmo: Cosmology | None = None, n_ell: int = 500, fwhm_arcmin: float = 1.4,
                  cmb_lmin: int = 8, cmb_lmax: int = 4000, window: bool = True, nells: int = 20,
                  ell_min: float = 30.0, ell_max: float = 3000.0) -> Problem:
    """The full 3x2pt layout with the CMB lensing probes of 6x2pt (w_gk,
    w_ks in the same 26 theta bins, C_kk at nells log-centred ells): a
    background that reaches z = 1300, a beam of fwhm_arcmin over
    [cmb_lmin, cmb_lmax] and optionally a pixel-window-like filter."""
    cosmo = cosmo or Cosmology()
    p = full_3x2pt_problem(cosmo, n_ell)
    p.chi_z, p.chi, p.G_z, p.G = background(cosmo, cmb=True)
    p.sixpt = 1
    p.cmb_fwhm, p.cmb_lmin, p.cmb_lmax = fwhm_arcmin, cmb_lmin, cmb_lmax
    p.cmb_win = healpix_window_like() if window else None
    p.kk_nells, p.kk_lmin, p.kk_lmax = nells, ell_min, ell_max
    return p


def kk_bandpower_binning(nbp: int = 8, lmin: int = 8, lmax: int = 400):
    """A synthetic kk band-power binning in init_cmb_auto_bandpower form:
    nbp top-hat bins in ln L over [lmin, lmax], each normalised to unit sum
    with an L(L+1) weight ramp, and a small theory offset per bin."""
    L = np.arange(lmin, lmax + 1)
    edges = np.unique(np.round(np.geomspace(lmin, lmax + 1, nbp + 1)).astype(int))
    M = np.zeros((nbp, len(L)))
    for j in range(nbp):
        sel = (L >= edges[j]) & (L < edges[j + 1])
        w = (L * (L + 1.0)) * sel
        M[j] = w / w.sum()
    off = 1e-9 * np.linspace(1.0, 2.0, nbp)
    return nbp, lmin, lmax, np.ascontiguousarray(M.reshape(-1)), off


def tatt_problem(cosmo: Cosmology | None = None, n_ell: int = 500, A2: float = -0.8, eta2: float = 0.5,
                 bta: float = 0.7) -> Problem:
    """The full 3x2pt layout with TATT intrinsic alignments (redshift
    evolution: A1 = 0.5, eta = 0; A2, eta2, b_TA as given)."""
    cosmo = cosmo or Cosmology()
    p = full_3x2pt_problem(cosmo, n_ell)
    p.ia_model, p.ia2_A, p.ia2_eta, p.ia_bta = 1, A2, eta2, bta
    p.ia_ps, p.ia_ps_kmin, p.ia_ps_kmax = fpt_ia_tables(cosmo)
    return p


def power_table(cosmo: Cosmology, nk: int = 400, nz: int = 100, kmin_log10=-4.0,
                kmax_log10=2.0, zmax=3.0):
    """(log10k[nk], z[nz], lnP[nk*nz] k-major) in (Mpc/h)^3, k in h/Mpc."""
    log10k = np.linspace(kmin_log10, kmax_log10, nk)
    z = np.linspace(0.0, zmax, nz)
    k = 10.0**log10k
    amp = (cosmo.sigma8 / sigma8_unnorm(cosmo.omega_m, cosmo.h, cosmo.ns)) ** 2
    p0 = amp * k**cosmo.ns * t_bbks(k, cosmo.omega_m, cosmo.h) ** 2
    d = growth_D(z, cosmo.omega_m)
    lnP = np.log(p0)[:, None] + 2.0 * np.log(d)[None, :]
    return log10k, z, np.ascontiguousarray(lnP.reshape(-1))


def linear_power_table(cosmo: Cosmology, nk: int = 300, nz: int = 60, kmin_log10=-4.0, kmax_log10=2.0,
                       zmax=4.0, k_nl=0.3):
    """Linear P(k,z) companion of ``power_table`` on its own grid: the
    ``power_table`` spectrum divided by a smooth nonlinear boost
    1 + (k D(z) / k_nl)^2, so P_lin < P_nl at high k and the two differ where
    the non-Limber correction (Limber with P_nl minus Limber with P_lin) reads
    them.  k-major ln P, (Mpc/h)^3, k in h/Mpc."""
    log10k = np.linspace(kmin_log10, kmax_log10, nk)
    z = np.linspace(0.0, zmax, nz)
    k = 10.0**log10k
    amp = (cosmo.sigma8 / sigma8_unnorm(cosmo.omega_m, cosmo.h, cosmo.ns)) ** 2
    p0 = amp * k**cosmo.ns * t_bbks(k, cosmo.omega_m, cosmo.h) ** 2
    d = growth_D(z, cosmo.omega_m)
    lnP = np.log(p0)[:, None] + 2.0 * np.log(d)[None, :] - np.log1p((k[:, None] * d[None, :] / k_nl) ** 2)
    return log10k, z, np.ascontiguousarray(lnP.reshape(-1))


def nonlimber_problem(cosmo: Cosmology | None = None, n_ell: int = 500, bmag=None) -> Problem:
    """The full 3x2pt layout with the non-Limber w(theta) below ell = 150
    (like.adopt_limber_gg = 0) and its linear P(k,z)."""
    cosmo = cosmo or Cosmology()
    p = full_3x2pt_problem(cosmo, n_ell)
    p.plin_log10k, p.plin_z, p.lnPL = linear_power_table(cosmo)
    p.limber_gg = 0
    if bmag is not None:
        p.bmag = np.asarray(bmag, dtype=np.float64)
    return p


def background(cosmo: Cosmology, cmb: bool = False):
    z = background_grid(cmb)
    chi = comoving_distance(z, cosmo.omega_m)
    d = growth_D(z, cosmo.omega_m)
    G = d * (1.0 + z)  # G = D/a
    return z, chi, z.copy(), G


# ----------------------------------------------------------------------------
# n(z) histograms (300 rows on [0, 3), dz = 0.01; column 0 = left edge)
# ----------------------------------------------------------------------------

def _zhisto(nrows=300, zmax=3.0):
    edges = np.linspace(0.0, zmax, nrows + 1)
    return edges[:-1], 0.5 * (edges[:-1] + edges[1:])


def source_table(nbin: int, z0: float, alpha=2.0, beta=1.5, sigma0=0.05, nrows=300,
                 zmax=3.0) -> np.ndarray:
    key = (nbin, z0, alpha, beta, sigma0, nrows, zmax)
    if key not in _NZ_CACHE:
        _NZ_CACHE[key] = _source_table(*key)
    return _NZ_CACHE[key].copy()


_NZ_CACHE: dict = {}


def _source_table(nbin, z0, alpha, beta, sigma0, nrows, zmax):
    """Equal-number photo-z bins of n(z) ~ z^alpha exp(-(z/z0)^beta).

    Photo-z scatter is Gaussian with sigma = sigma0 (1+z).  Returns a
    (nrows, nbin+1) table: z_left, n_1(z), ..., n_nbin(z).
    """
    zl, zc = _zhisto(nrows, zmax)
    zf = np.linspace(0.0, zmax, 6001)
    nf = zf**alpha * np.exp(-((zf / z0) ** beta))
    # photo-z distribution of the full sample, for equal-number edges
    zp = np.linspace(-0.5, zmax + 1.0, 4001)
    sig = sigma0 * (1.0 + zf)
    kern = np.exp(-0.5 * ((zp[:, None] - zf[None, :]) / sig[None, :]) ** 2) / sig[None, :]
    nph = integrate.simpson(kern * nf[None, :], x=zf, axis=1)
    cph = integrate.cumulative_simpson(nph, x=zp, initial=0.0)
    cph /= cph[-1]
    edges = np.interp(np.linspace(0.0, 1.0, nbin + 1), cph, zp)
    edges[0], edges[-1] = -np.inf, np.inf
    nzc = zc**alpha * np.exp(-((zc / z0) ** beta))
    s = sigma0 * (1.0 + zc)
    out = np.empty((nrows, nbin + 1))
    out[:, 0] = zl
    for i in range(nbin):
        lo = special.erf((edges[i] - zc) / (np.sqrt(2.0) * s)) if np.isfinite(edges[i]) else -1.0
        hi = special.erf((edges[i + 1] - zc) / (np.sqrt(2.0) * s)) if np.isfinite(edges[i + 1]) else 1.0
        out[:, i + 1] = nzc * 0.5 * (hi - lo)
    return out


def lens_table(centres, sigma=0.05, nrows=300, zmax=3.0) -> np.ndarray:
    zl, zc = _zhisto(nrows, zmax)
    out = np.empty((nrows, len(centres) + 1))
    out[:, 0] = zl
    for i, c in enumerate(centres):
        out[:, i + 1] = np.exp(-0.5 * ((zc - c) / sigma) ** 2)
    return out


# ----------------------------------------------------------------------------
# Full problem descriptions
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class Problem:
    """Everything one data vector needs (one draw)."""
    omega_m: float
    chi_z: np.ndarray
    chi: np.ndarray
    G_z: np.ndarray
    G: np.ndarray
    log10k: np.ndarray
    pk_z: np.ndarray
    lnP: np.ndarray
    source: np.ndarray           # (nrows, ns+1)
    lens: np.ndarray | None      # (nrows, nl+1)
    ntheta: int
    theta_min_arcmin: float
    theta_max_arcmin: float
    n_ell: int = 300
    lmax: int = 50000
    hdi: int = 0                 # Ntable.high_def_integration
    src_dz: np.ndarray | None = None
    lens_dz: np.ndarray | None = None
    lens_stretch: np.ndarray | None = None
    b1: np.ndarray | None = None
    bmag: np.ndarray | None = None
    shear_m: np.ndarray | None = None
    ia_A: float = 0.5
    ia_eta: float = 0.0
    ia_mode: int = 3             # IA_REDSHIFT_EVOLUTION (cosmolike/IA.h:13)
    h: float = 0.7
    pm: np.ndarray | None = None  # point-mass amplitudes per lens bin
    # TATT intrinsic alignments (IA_MODEL = 1): A2 (value, exponent), b_TA and
    # the FAST-PT IA tables in set_IA_PS form ([12][N], k range in h/Mpc)
    ia_model: int = 0
    ia2_A: float = 0.0
    ia2_eta: float = 0.0
    ia_bta: float = 0.0
    ia_ps: np.ndarray | None = None
    ia_ps_kmin: float = 0.0
    ia_ps_kmax: float = 0.0
    # one-loop galaxy bias: b2 (bs2 follows from b1), b3, bK per lens bin and
    # the FAST-PT bias tables in set_bias_PS form ([12][N], k range in h/Mpc)
    b2: np.ndarray | None = None
    b3: np.ndarray | None = None
    bk: np.ndarray | None = None
    bias_ps: np.ndarray | None = None
    bias_ps_kmin: float = 0.0
    bias_ps_kmax: float = 0.0
    bias_ps_sigma4: float = 0.0
    # 6x2pt: CMB lensing probes, the CMB filter and the kk ell bins
    sixpt: int = 0
    cmb_fwhm: float | None = None
    cmb_lmin: int = 0
    cmb_lmax: int = 0
    cmb_win: np.ndarray | None = None
    kk_nells: int = 0
    kk_lmin: float = 0.0
    kk_lmax: float = 0.0
    kk_bp: tuple | None = None   # (nbp, lmin, lmax, M [nbp*(lmax-lmin+1)], offset [nbp]) band powers
    # non-Limber w (like.adopt_limber_gg = 0) and its linear P(k,z)
    limber_gg: int = 1
    plin_log10k: np.ndarray | None = None
    plin_z: np.ndarray | None = None
    lnPL: np.ndarray | None = None

    @property
    def ns(self):
        return self.source.shape[1] - 1

    @property
    def nl(self):
        return 0 if self.lens is None else self.lens.shape[1] - 1

    def fill_defaults(self):
        ns, nl = self.ns, max(self.nl, 1)
        if self.src_dz is None:
            self.src_dz = np.zeros(ns)
        if self.shear_m is None:
            self.shear_m = np.zeros(ns)
        if self.lens_dz is None:
            self.lens_dz = np.zeros(nl)
        if self.lens_stretch is None:
            self.lens_stretch = np.ones(nl)
        if self.b1 is None:
            self.b1 = np.linspace(1.2, 2.0, nl) if nl > 1 else np.array([1.2])
        if self.bmag is None:
            self.bmag = np.zeros(nl)
        if self.pm is None:
            self.pm = np.zeros(nl)
        return self


_BG_CACHE: dict = {}


def _cosmo_tables(cosmo: Cosmology):
    key = (cosmo.omega_m, cosmo.h, cosmo.ns, cosmo.sigma8)
    if key not in _BG_CACHE:
        z, chi, gz, G = background(cosmo)
        log10k, pz, lnP = power_table(cosmo)
        _BG_CACHE[key] = (z, chi, gz, G, log10k, pz, lnP)
    return _BG_CACHE[key]


def shear_only_problem(cosmo: Cosmology | None = None) -> Problem:
    """Cosmic shear, 5 source bins, 20 theta bins 2.5'-250', N_ell=300."""
    cosmo = cosmo or Cosmology()
    z, chi, gz, G, log10k, pz, lnP = _cosmo_tables(cosmo)
    src = source_table(5, 0.5)
    return Problem(cosmo.omega_m, z, chi, gz, G, log10k, pz, lnP, src, None, 20, 2.5,
                   250.0, n_ell=300, h=cosmo.h).fill_defaults()


def full_3x2pt_problem(cosmo: Cosmology | None = None, n_ell: int = 500) -> Problem:
    """10 Smail source bins (z0=0.6), 5 Gaussian lens bins, 26 theta bins 2.5'-900'."""
    cosmo = cosmo or Cosmology()
    z, chi, gz, G, log10k, pz, lnP = _cosmo_tables(cosmo)
    src = source_table(10, 0.6)
    lens = lens_table(np.linspace(0.3, 0.9, 5), 0.05)
    p = Problem(cosmo.omega_m, z, chi, gz, G, log10k, pz, lnP, src, lens, 26, 2.5, 900.0,
                n_ell=n_ell, h=cosmo.h)
    p.b1 = np.linspace(1.2, 2.0, 5)
    return p.fill_defaults()


def ragged_3x2pt_problem(cosmo: Cosmology, ns: int, nl: int, ntheta: int, n_ell: int = 500) -> Problem:
    """3x2pt layout with ns Smail source bins (z0=0.6), nl Gaussian lens bins
    centred 0.3-0.9 (sigma_z=0.05) and ntheta theta bins 2.5'-900'."""
    z, chi, gz, G, log10k, pz, lnP = _cosmo_tables(cosmo)
    src = source_table(ns, 0.6)
    lens = lens_table(np.linspace(0.3, 0.9, nl), 0.05)
    p = Problem(cosmo.omega_m, z, chi, gz, G, log10k, pz, lnP, src, lens, ntheta, 2.5, 900.0,
                n_ell=n_ell, h=cosmo.h)
    p.b1 = np.linspace(1.2, 2.0, nl)
    return p.fill_defaults()


def random_draws(n: int, seed: int = 20261203, ragged: bool = True):
    """Cosmology/nuisance draws; ragged layouts when ``ragged``."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        c = Cosmology(omega_m=rng.uniform(0.2, 0.4), h=rng.uniform(0.6, 0.8),
                      sigma8=rng.uniform(0.6, 1.0))
        if ragged:
            ns, nl, nt = int(rng.integers(3, 11)), int(rng.integers(2, 7)), int(rng.integers(20, 31))
        else:
            ns, nl, nt = 10, 5, 26
        out.append((c, ns, nl, nt))
    return out
