"""Python mirror of the cosmolike_interface calls Cocoa makes for the
real-space 3x2pt path (cosmolike/generic_interface.{hpp,cpp},
cosmolike/cosmo2D_wrapper.cpp), backed by the HIP library.

Method names follow the reference; errors raise ``CosmolikeError`` (the
reference logs and exits).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import D, clh_draw


class CosmolikeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_lib.ERRORS.get(code, code)}: {msg}")
        self.code = code


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _p(x):
    return None if x is None else x.ctypes.data_as(D)


LEGENDRE_FOLDED = 0
LEGENDRE_FULL = 1
PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2}


class Layout:
    """A captured survey layout (owns a clh_layout handle)."""

    def __init__(self, lib, handle):
        self.L = lib
        self.h = handle

    def datavector_size(self):
        return self.L.clh_layout_datavector_size(self.h)

    def __del__(self):
        try:
            if self.h:
                self.L.clh_layout_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Cosmolike:
    """One library context: survey layout + the current draw."""

    def __init__(self, device: int = 0, handle=None):
        self.L = _lib.lib()
        if handle is not None:
            self.h = ctypes.c_void_p(handle)
            self._owned = False
        else:
            h = ctypes.c_void_p()
            rc = self.L.clh_create(ctypes.byref(h), device)
            if rc:
                raise CosmolikeError(rc, "clh_create failed")
            self.h = h
            self._owned = True
        self.ns = self.nl = self.ntheta = 0
        self._keep = []

    @classmethod
    def default(cls):
        """The process-wide context behind xi_pm_tomo & co."""
        L = _lib.lib()
        return cls(handle=L.clh_default())

    def close(self):
        if getattr(self, "_owned", False) and self.h:
            self.L.clh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise CosmolikeError(rc, self.L.clh_last_error(self.h).decode())

    # ------------------------------------------------------------- init/set
    def init_ntable(self, n_ell=300, n_a=225, lmax=50000, high_def=0):
        self._chk(self.L.clh_init_ntable(self.h, n_ell, n_a, lmax, high_def))

    def init_binning_real_space(self, ntheta, tmin_arcmin, tmax_arcmin):
        self.ntheta = ntheta
        self._chk(self.L.clh_init_binning_real_space(self.h, ntheta, tmin_arcmin, tmax_arcmin))

    def init_IA(self, ia_model=0, ia_redshift_evol=3):
        self._chk(self.L.clh_init_ia(self.h, ia_model, ia_redshift_evol))

    def set_legendre_mode(self, mode):
        self._chk(self.L.clh_set_legendre_mode(self.h, mode))

    def set_precision(self, mode):
        """``"fp64"`` (plain path), ``"fp32"`` or ``"bf16x3"``, or a CLH_PRECISION_* value."""
        if isinstance(mode, str):
            mode = PRECISION[mode]
        self._chk(self.L.clh_set_precision(self.h, int(mode)))

    def set_cosmological_parameters(self, omega_matter, hubble=70.0):
        self._chk(self.L.clh_set_cosmological_parameters(self.h, float(omega_matter), float(hubble)))

    def set_distances(self, z, chi):
        z, chi = _arr(z), _arr(chi)
        self._chk(self.L.clh_set_distances(self.h, _p(z), _p(chi), len(z)))

    def set_growth(self, z, G):
        z, G = _arr(z), _arr(G)
        self._chk(self.L.clh_set_growth(self.h, _p(z), _p(G), len(z)))

    def set_non_linear_power_spectrum(self, log10k, z, lnP):
        lk, z, lnP = _arr(log10k), _arr(z), _arr(lnP)
        self._chk(self.L.clh_set_non_linear_power_spectrum(self.h, _p(lk), len(lk), _p(z), len(z), _p(lnP)))

    def set_source_sample(self, table):
        t = _arr(table)
        self.ns = t.shape[1] - 1
        self._chk(self.L.clh_set_source_sample(self.h, _p(t), t.shape[0], self.ns))

    def set_lens_sample(self, table):
        t = _arr(table)
        self.nl = t.shape[1] - 1
        self._chk(self.L.clh_set_lens_sample(self.h, _p(t), t.shape[0], self.nl))

    def _setv(self, fn, v):
        v = _arr(v)
        self._chk(fn(self.h, _p(v), len(v)))

    def set_nuisance_shear_calib(self, m):
        self._setv(self.L.clh_set_nuisance_shear_calib, m)

    def set_nuisance_shear_photoz(self, dz):
        self._setv(self.L.clh_set_nuisance_shear_photoz, dz)

    def set_nuisance_clustering_photoz(self, dz):
        self._setv(self.L.clh_set_nuisance_clustering_photoz, dz)

    def set_nuisance_clustering_photoz_stretch(self, s):
        self._setv(self.L.clh_set_nuisance_clustering_photoz_stretch, s)

    def set_nuisance_linear_bias(self, b1):
        self._setv(self.L.clh_set_nuisance_linear_bias, b1)

    def set_nuisance_magnification_bias(self, bmag):
        self._setv(self.L.clh_set_nuisance_magnification_bias, bmag)

    def set_nuisance_IA(self, A1, A2=None, BTA=None):
        """set_nuisance_IA: A1 only (NLA), or A1, A2 and b_TA (TATT); mode 2
        one value per source bin in each, mode 3 (A, eta), (A2, eta2), (b_TA,)."""
        if A2 is None and BTA is None:
            self._setv(self.L.clh_set_nuisance_ia, A1)
            return
        A1, A2, BTA = _arr(A1), _arr(A2), _arr(BTA)
        n = min(len(A1), len(A2))
        BTA = np.resize(BTA, n)
        self._chk(self.L.clh_set_nuisance_ia_tatt(self.h, _p(A1), _p(A2), _p(BTA), n))

    def set_nuisance_nonlinear_bias(self, b1, b2):
        """b2 per lens bin (bs2 = -4/7 (b1 - 1) where b2 != 0)."""
        b1, b2 = _arr(b1), _arr(b2)
        self._chk(self.L.clh_set_nuisance_nonlinear_bias(self.h, _p(b1), _p(b2), len(b2)))

    def set_nuisance_nonlocal_bias(self, b3, bk):
        b3, bk = _arr(b3), _arr(bk)
        self._chk(self.L.clh_set_nuisance_nonlocal_bias(self.h, _p(b3), _p(bk), len(b3)))

    def set_nuisance_bias_fastpt(self, b1, b2, bmag, b3, bk):
        """set_nuisance_bias_fastpt: linear, nonlinear, magnification and nonlocal bias."""
        self.set_nuisance_linear_bias(b1)
        self.set_nuisance_nonlinear_bias(b1, b2)
        self.set_nuisance_magnification_bias(bmag)
        self.set_nuisance_nonlocal_bias(b3, bk)

    def set_bias_PS(self, PS, kmin, kmax, sigma4, cutoff=1.0e4):
        """set_bias_PS: the FAST-PT bias spectra PS[12][N] on N log-spaced k in [kmin, kmax) h/Mpc."""
        PS = _arr(PS)
        n = PS.size // 12
        if PS.size != 12 * n:
            raise ValueError("PS must hold 12 rows of N entries")
        self._chk(self.L.clh_set_bias_PS(self.h, _p(PS), float(kmin), float(kmax), float(cutoff), float(sigma4), n))

    def init_probes(self, probes="3x2pt"):
        """init_probes: "3x2pt" or "6x2pt" (adds w_gk, w_ks and C_kk blocks)."""
        self._chk(self.L.clh_init_probes(self.h, probes.encode()))

    def init_cmb_cross_correlation(self, lmin, lmax, fwhm_arcmin, healpix_window=None):
        w = None if healpix_window is None else _arr(healpix_window)
        self._chk(self.L.clh_init_cmb_cross_correlation(self.h, int(lmin), int(lmax), float(fwhm_arcmin), _p(w),
                                                        0 if w is None else len(w)))

    def init_cmb_auto_bandpower(self, nbins, lmin, lmax, binning_matrix, theory_offset=None, alpha=1.0):
        """init_cmb_auto_bandpower: the kk block as nbins band powers; the
        binning matrix [nbins][lmax - lmin + 1] and offset [nbins] as arrays."""
        M = _arr(binning_matrix)
        if M.size != int(nbins) * (int(lmax) - int(lmin) + 1):
            raise ValueError("binning matrix must hold nbins x (lmax - lmin + 1) entries")
        off = None if theory_offset is None else _arr(theory_offset)
        if off is not None and off.size != int(nbins):
            raise ValueError("theory offset must hold nbins entries")
        self._chk(self.L.clh_init_cmb_auto_bandpower(self.h, int(nbins), int(lmin), int(lmax), _p(M), _p(off),
                                                     float(alpha)))

    def kk_bandpower_info(self):
        """(nbins, lmin, lmax, alpha) of the kk band powers (nbins 0: none)."""
        n, lo, hi = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        a = np.zeros(1)
        self._chk(self.L.clh_cmb_kk_bandpower_info(self.h, ctypes.byref(n), ctypes.byref(lo), ctypes.byref(hi),
                                                   _p(a)))
        return n.value, lo.value, hi.value, float(a[0])

    def compute_6x2pt_real(self, ord=(0, 1, 2, 3, 4, 5)):
        """compute_Mx2pt_N_masked<0,6>(ord)."""
        n = self.datavector_size()
        out = np.empty(n)
        o = (ctypes.c_int * 6)(*ord)
        self._chk(self.L.clh_compute_6x2pt_real(self.h, o, _p(out), n))
        return out

    def set_linear_power_spectrum(self, log10k, z, lnPL):
        lk, z, lnPL = _arr(log10k), _arr(z), _arr(lnPL)
        self._chk(self.L.clh_set_linear_power_spectrum(self.h, _p(lk), len(lk), _p(z), len(z), _p(lnPL)))

    def set_adopt_limber_gg(self, limber):
        """like.adopt_limber_gg: 1 Limber w(theta); 0 the non-Limber C_gg below ell 150."""
        self._chk(self.L.clh_set_adopt_limber_gg(self.h, int(limber)))

    def cl_nonlimber(self):
        """C_cl_tomo as w uses it: (cl[nl][150] at ell 0..149, LMAX[nl])."""
        cl = np.empty((self.nl, 150))
        lmax = np.empty(self.nl, dtype=np.int32)
        self._chk(self.L.clh_cl_nonlimber(self.h, _p(cl), lmax.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return cl, lmax

    def set_IA_PS(self, PS, kmin, kmax, cutoff=1.0e4):
        """set_IA_PS: the FAST-PT IA spectra PS[12][N] on N log-spaced k in
        [kmin, kmax) h/Mpc."""
        PS = _arr(PS)
        n = PS.size // 12
        if PS.size != 12 * n:
            raise ValueError("PS must hold 12 rows of N entries")
        self._chk(self.L.clh_set_IA_PS(self.h, _p(PS), float(kmin), float(kmax), float(cutoff), n))

    def load(self, prob, legendre_mode=LEGENDRE_FOLDED):
        """Configure the layout and current draw from a synth.Problem."""
        self.init_ntable(prob.n_ell, 225, prob.lmax, getattr(prob, "hdi", 0))
        self.init_binning_real_space(prob.ntheta, prob.theta_min_arcmin, prob.theta_max_arcmin)
        self.init_IA(getattr(prob, "ia_model", 0), prob.ia_mode)
        self.set_legendre_mode(legendre_mode)
        self.set_adopt_limber_gg(getattr(prob, "limber_gg", 1))
        if getattr(prob, "sixpt", 0):
            self.init_probes("6x2pt")
            self.init_cmb_cross_correlation(prob.cmb_lmin, prob.cmb_lmax, prob.cmb_fwhm, prob.cmb_win)
            self.init_binning_fourier(prob.kk_nells, int(prob.kk_lmin), int(prob.kk_lmax), int(prob.kk_lmax))
            if getattr(prob, "kk_bp", None) is not None:
                nbp, lo, hi, M, off = prob.kk_bp
                self.init_cmb_auto_bandpower(nbp, lo, hi, M, off)
        else:
            self.init_probes("3x2pt")
        self.set_source_sample(prob.source)
        if prob.lens is not None:
            self.set_lens_sample(prob.lens)
        self.init_ggl_exclude(getattr(prob, "ggl_exclude", None))
        self.set_draw(prob)
        return self

    def set_draw(self, prob):
        self.set_cosmological_parameters(prob.omega_m, 100.0 * prob.h)
        self.set_distances(prob.chi_z, prob.chi)
        self.set_growth(prob.G_z, prob.G)
        self.set_non_linear_power_spectrum(prob.log10k, prob.pk_z, prob.lnP)
        if getattr(prob, "lnPL", None) is not None:
            self.set_linear_power_spectrum(prob.plin_log10k, prob.plin_z, prob.lnPL)
        self.set_nuisance_shear_photoz(prob.src_dz)
        self.set_nuisance_shear_calib(prob.shear_m)
        if prob.lens is not None:
            self.set_nuisance_clustering_photoz(prob.lens_dz)
            self.set_nuisance_clustering_photoz_stretch(prob.lens_stretch)
            self.set_nuisance_linear_bias(prob.b1)
            self.set_nuisance_magnification_bias(prob.bmag)
            if getattr(prob, "pm", None) is not None:
                self.set_nuisance_pointmass(prob.pm)
            if getattr(prob, "b2", None) is not None:
                self.set_bias_PS(prob.bias_ps, prob.bias_ps_kmin, prob.bias_ps_kmax, prob.bias_ps_sigma4)
                self.set_nuisance_nonlinear_bias(prob.b1, prob.b2)
                self.set_nuisance_nonlocal_bias(prob.b3, prob.bk)
            else:
                z = np.zeros(prob.nl)
                self.set_nuisance_nonlinear_bias(prob.b1, z)
                self.set_nuisance_nonlocal_bias(z, z)
        tatt = getattr(prob, "ia_model", 0) == 1
        if tatt:
            self.set_IA_PS(prob.ia_ps, prob.ia_ps_kmin, prob.ia_ps_kmax)
        if prob.ia_mode == 3 and tatt:
            self.set_nuisance_IA([prob.ia_A, prob.ia_eta], [prob.ia2_A, prob.ia2_eta], [prob.ia_bta])
        elif prob.ia_mode == 3:
            self.set_nuisance_IA([prob.ia_A, prob.ia_eta])

    def init_binning_fourier(self, nells, lmin, lmax, lmax_shear):
        self._chk(self.L.clh_init_binning_fourier(self.h, int(nells), int(lmin), int(lmax), int(lmax_shear)))

    def datavector_size_fourier(self):
        return self.L.clh_datavector_size_fourier(self.h)

    def compute_3x2pt_fourier(self, ord=(0, 1, 2), mask=None):
        """compute_Mx2pt_N_masked<1,3>(ord); mask in the order of the output
        (None = all ones)."""
        n = self.L.clh_datavector_size_fourier(self.h)
        out = np.empty(n)
        m = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask, dtype=np.int32))
            if m.size != n:
                raise ValueError(f"mask has {m.size} entries, data vector {n}")
        I3 = (ctypes.c_int * 3)(*ord)
        self._chk(self.L.clh_compute_3x2pt_fourier(
            self.h, I3, None if m is None else m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(out), n))
        return out

    def init_ggl_exclude(self, pairs):
        """init_ggl_exclude: (lens, source) pairs left out of gamma_t, flat
        [zl0, zs0, zl1, zs1, ...] or a list of pairs; None or empty clears."""
        a = np.ascontiguousarray(np.asarray([] if pairs is None else pairs, dtype=np.int32).ravel())
        self._chk(self.L.clh_init_ggl_exclude(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
                                              if a.size else None, int(a.size)))

    def test_zoverlap(self, zl, zs):
        r = self.L.clh_test_zoverlap(self.h, int(zl), int(zs))
        if r < 0:
            self._chk(r)
        return r

    def int_for(self, probe, params, a):
        """Limber integrand of probe 0 ss .. 5 kk at the scale factors a, with the
        reference's parameter array (clh_int_for_limber)."""
        a = _arr(a)
        ar = _arr(params)
        out = np.empty_like(a)
        self._chk(self.L.clh_int_for_limber(self.h, int(probe), _p(ar), _p(a), len(a), _p(out)))
        return out

    @property
    def n_ggl(self):
        """Included gs pairs (tomo.ggl_Npowerspectra)."""
        return sum(self.test_zoverlap(zl, zs) for zl in range(self.nl) for zs in range(self.ns))

    def set_nuisance_pointmass(self, pm):
        pm = _arr(pm)
        self._chk(self.L.clh_set_nuisance_pointmass(self.h, _p(pm), len(pm)))

    # ----------------------------------------------------------- likelihood
    @staticmethod
    def _ord(ord):
        return (ctypes.c_int * 3)(*ord)

    def like_set_mask(self, mask, ord=(0, 1, 2)):
        """IP::set_mask: 0/1 per data-vector entry in the block order ord (None clears)."""
        if mask is None:
            self._chk(self.L.clh_like_set_mask(self.h, None, 0, self._ord(ord)))
            return
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.int32))
        self._chk(self.L.clh_like_set_mask(self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(m),
                                           self._ord(ord)))

    def like_set_data(self, data, ord=(0, 1, 2)):
        d = _arr(data)
        self._chk(self.L.clh_like_set_data(self.h, _p(d), len(d), self._ord(ord)))

    def like_set_inv_cov(self, inv_cov, ord=(0, 1, 2)):
        c = _arr(inv_cov)
        self._chk(self.L.clh_like_set_inv_cov(self.h, _p(c), c.shape[0], self._ord(ord)))

    def chi2(self, dv):
        """IP::get_chi2 of one (host) data vector."""
        d = _arr(dv)
        out = ctypes.c_double()
        self._chk(self.L.clh_chi2(self.h, _p(d), len(d), ctypes.byref(out)))
        return out.value

    def batch_chi2(self, dv_ptr, chi2_ptr, grad_ptr=0, ord=(0, 1, 2)):
        """Device pointers: dv [B][n] in, chi2 [B] and d chi2/d dv [B][n] (0 = skip) out."""
        self._chk(self.L.clh_batch_chi2(self.h, self._ord(ord), ctypes.c_void_p(dv_ptr), ctypes.c_void_p(chi2_ptr),
                                        ctypes.c_void_p(grad_ptr or None)))

    # -------------------------------------------------------------- compute
    @property
    def n_ss(self):
        return self.ns * (self.ns + 1) // 2

    @property
    def n_gs(self):
        return self.ns * self.nl

    def datavector_size(self):
        return self.L.clh_datavector_size(self.h)

    def compute_3x2pt_real(self, ord=(0, 1, 2)):
        """compute_Mx2pt_N_masked<0,3>(ord) with an all-ones mask."""
        n = self.datavector_size()
        out = np.empty(n)
        o = (ctypes.c_int * 3)(*ord)
        self._chk(self.L.clh_compute_3x2pt_real(self.h, o, _p(out), n))
        return out

    def xi_pm(self):
        xp = np.empty(self.n_ss * self.ntheta)
        xm = np.empty(self.n_ss * self.ntheta)
        self._chk(self.L.clh_xi_pm(self.h, _p(xp), _p(xm)))
        return xp.reshape(self.n_ss, self.ntheta), xm.reshape(self.n_ss, self.ntheta)

    def gammat(self):
        """[included pairs][ntheta] (every pair without init_ggl_exclude)."""
        n = self.n_ggl
        out = np.empty(n * self.ntheta)
        self._chk(self.L.clh_gammat(self.h, _p(out)))
        return out.reshape(n, self.ntheta)

    def wtheta(self):
        out = np.empty(self.nl * self.ntheta)
        self._chk(self.L.clh_wtheta(self.h, _p(out)))
        return out.reshape(self.nl, self.ntheta)

    def cl_tables(self, n_ell):
        ss = np.empty((self.n_ss, n_ell))
        gs = np.empty((max(self.n_gs, 1), n_ell))
        gg = np.empty((max(self.nl, 1), n_ell))
        self._chk(self.L.clh_cl_tables(self.h, _p(ss), _p(gs), _p(gg)))
        return ss, gs[: self.n_gs], gg[: self.nl]

    def cl_tables_bb(self, n_ell):
        """B-mode C_ss table [n_ss][N_ell] (zero under NLA)."""
        bb = np.empty((self.n_ss, n_ell))
        self._chk(self.L.clh_cl_tables_bb(self.h, _p(bb)))
        return bb

    def cl_nointerp(self, probe, ells):
        """probe "ss" (E mode), "gs", "gg", "ss_bb" (B mode), and with the 6x2pt
        probes "gk", "ks", "kk"."""
        ells = _arr(ells)
        idx = {"ss": 0, "gs": 1, "gg": 2, "ss_bb": 3, "gk": 4, "ks": 5, "kk": 6}[probe]
        n = {"ss": self.n_ss, "gs": self.n_gs, "gg": self.nl, "ss_bb": self.n_ss, "gk": self.nl, "ks": self.ns,
             "kk": 1}[probe]
        out = np.empty((n, len(ells)))
        self._chk(self.L.clh_cl_nointerp(self.h, idx, _p(ells), len(ells), _p(out)))
        return out

    # ------------------------------------------------------------- batches
    @staticmethod
    def _draw_array(problems):
        arr = (clh_draw * len(problems))()
        keep = []
        for i, p in enumerate(problems):
            vals = {k: _arr(getattr(p, k)) for k in ("chi_z", "chi", "G_z", "G", "log10k", "pk_z", "lnP",
                                                     "src_dz", "shear_m")}
            for k in ("lens_dz", "lens_stretch", "b1", "bmag", "pm"):
                v = getattr(p, k, None)
                vals[k] = _arr(v) if (p.lens is not None and v is not None) else None
            keep.append(vals)
            d = arr[i]
            d.omega_m = p.omega_m
            d.chi_z, d.chi, d.chi_n = _p(vals["chi_z"]), _p(vals["chi"]), len(vals["chi_z"])
            d.G_z, d.G, d.G_n = _p(vals["G_z"]), _p(vals["G"]), len(vals["G_z"])
            d.log10k, d.pk_z, d.lnP = _p(vals["log10k"]), _p(vals["pk_z"]), _p(vals["lnP"])
            d.nk, d.npz = len(vals["log10k"]), len(vals["pk_z"])
            d.src_dz, d.shear_m = _p(vals["src_dz"]), _p(vals["shear_m"])
            d.lens_dz, d.lens_stretch = _p(vals["lens_dz"]), _p(vals["lens_stretch"])
            d.b1, d.bmag = _p(vals["b1"]), _p(vals["bmag"])
            d.pm = _p(vals["pm"])
            d.ia_A, d.ia_eta = p.ia_A, p.ia_eta
            d.h = getattr(p, "h", 0.7)
            if p.lens is not None and getattr(p, "b2", None) is not None:
                for k in ("b2", "b3", "bk", "bias_ps"):
                    vals[k] = _arr(getattr(p, k))
                d.b2, d.b3, d.bk = _p(vals["b2"]), _p(vals["b3"]), _p(vals["bk"])
                d.fptb, d.fptb_n = _p(vals["bias_ps"]), vals["bias_ps"].size // 12
                d.fptb_kmin, d.fptb_kmax, d.fptb_sigma4 = p.bias_ps_kmin, p.bias_ps_kmax, p.bias_ps_sigma4
            if getattr(p, "lnPL", None) is not None:
                for k in ("plin_log10k", "plin_z", "lnPL"):
                    vals[k] = _arr(getattr(p, k))
                d.plin_log10k, d.plin_z, d.lnPL = _p(vals["plin_log10k"]), _p(vals["plin_z"]), _p(vals["lnPL"])
                d.plin_nk, d.plin_nz = len(vals["plin_log10k"]), len(vals["plin_z"])
            if getattr(p, "ia_model", 0) == 1:
                vals["fpt"] = _arr(p.ia_ps)
                d.ia2_A, d.ia2_eta, d.bta_A = p.ia2_A, p.ia2_eta, p.ia_bta
                d.fpt, d.fpt_n = _p(vals["fpt"]), vals["fpt"].size // 12
                d.fpt_kmin, d.fpt_kmax = p.ia_ps_kmin, p.ia_ps_kmax
        return arr, keep

    def batch_upload(self, problems):
        """Upload a list of synth.Problem draws (same layout as the context)."""
        arr, keep = self._draw_array(problems)
        self._chk(self.L.clh_batch_upload(self.h, len(problems), arr))
        self.batch_size = len(problems)
        self.param_generation = getattr(self, "param_generation", 0) + 1
        # the uploaded values of the differentiable inputs, [B][...] in the
        # order of batch_set_params (None where the layout has no such input)
        lens = problems[0].lens is not None
        self._uploaded = [
            np.stack([_arr(p.lnP).ravel() for p in problems]),
            np.stack([_arr(p.src_dz) for p in problems]),
            np.stack([_arr(p.lens_dz) for p in problems]) if lens else None,
            np.stack([_arr(p.b1) for p in problems]) if lens else None,
            np.array([float(p.ia_A) for p in problems]),
            np.stack([_arr(p.shear_m) for p in problems]),
        ]
        self._uploaded_dev = None

    def uploaded_params(self, device):
        """The uploaded batch's (lnP, src_dz, lens_dz, b1, ia_A, shear_m) as
        float64 tensors on `device` (None where the layout has no such input)."""
        import torch
        key = (self.param_generation, str(device))
        if getattr(self, "_uploaded_dev", None) is None or self._uploaded_dev[0] != key:
            t = [None if a is None else torch.from_numpy(a).to(device) for a in self._uploaded]
            self._uploaded_dev = (key, t)
        return self._uploaded_dev[1]

    def capture_layout(self):
        """Snapshot of the current survey layout (clh_layout_capture)."""
        h = ctypes.c_void_p()
        self._chk(self.L.clh_layout_capture(self.h, ctypes.byref(h)))
        return Layout(self.L, h)

    def batch_upload_layouts(self, problems, layouts):
        """Upload draws with per-draw captured layouts (forward only)."""
        arr, keep = self._draw_array(problems)
        hs = (ctypes.c_void_p * len(layouts))(*[l.h.value for l in layouts])
        self._chk(self.L.clh_batch_upload_layouts(self.h, len(problems), arr, hs))
        self.batch_size = len(problems)

    def batch_offsets(self):
        """Data-vector offsets of the uploaded draws (length B+1)."""
        return [int(self.L.clh_batch_datavector_offset(self.h, i)) for i in range(self.batch_size + 1)]

    def batch_forward(self, dv_device_ptr: int, ord=(0, 1, 2)):
        o = (ctypes.c_int * 3)(*ord)
        self._chk(self.L.clh_batch_forward(self.h, o, ctypes.c_void_p(dv_device_ptr)))

    def batch_use_graph(self, on=True):
        self._chk(self.L.clh_batch_use_graph(self.h, 1 if on else 0))

    def set_stream(self, stream_ptr):
        self._chk(self.L.clh_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def batch_set_params(self, lnP=0, src_dz=0, lens_dz=0, b1=0, ia_A=0, shear_m=0):
        """Device pointers (ints; 0 = keep) of [B][...] parameter arrays."""
        v = [ctypes.c_void_p(x or None) for x in (lnP, src_dz, lens_dz, b1, ia_A, shear_m)]
        self._chk(self.L.clh_batch_set_params(self.h, *v))
        self.param_generation = getattr(self, "param_generation", 0) + 1

    def batch_vjp(self, dv_bar, g_lnP=0, g_src_dz=0, g_lens_dz=0, g_b1=0, g_ia_A=0, g_m=0, ord=(0, 1, 2)):
        """Device pointers (ints; 0 = skip) of the cotangent and the gradient outputs."""
        v = [ctypes.c_void_p(x or None) for x in (dv_bar, g_lnP, g_src_dz, g_lens_dz, g_b1, g_ia_A, g_m)]
        o = (ctypes.c_int * 3)(*ord)
        self._chk(self.L.clh_batch_vjp(self.h, o, *v))

    def synchronize(self):
        self._chk(self.L.clh_synchronize(self.h))

    def autotune(self):
        """Pick the rows per tile of the Limber stages (and LDS staging of the
        gg tables) by timing the uploaded batch; returns the rows per tile."""
        rows = (ctypes.c_int * 4)()
        lds = (ctypes.c_int * 2)()
        self._chk(self.L.clh_autotune(self.h, rows, lds))
        self.gg_lds = list(lds)
        return list(rows)

    def set_tile_rows(self, rows):
        r = (ctypes.c_int * 4)(*rows)
        self._chk(self.L.clh_set_tile_rows(self.h, r))

    def tile_rows(self):
        rows = (ctypes.c_int * 4)()
        self._chk(self.L.clh_get_tile_rows(self.h, rows))
        return list(rows)

    def set_vjp_checkpoint(self, on=True):
        """Batch forwards keep the gg integrand for a following VJP."""
        self._chk(self.L.clh_set_vjp_checkpoint(self.h, 1 if on else 0))

    def set_vjp_lens_limits(self, mode=0):
        """Lens-limit part of the lens-shift gradient: 0 from the nodes'
        motion inside the adjoint kernels (default), 1 by finite-difference
        views (a cross-check), 2 left out (the gradient at fixed limits)."""
        self._chk(self.L.clh_set_vjp_lens_limits(self.h, int(mode)))

    def enable_timing(self, on=True):
        self._chk(self.L.clh_enable_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        names = (ctypes.c_char_p * 9)()
        ms = (ctypes.c_float * 9)()
        n = self.L.clh_last_kernel_times(self.h, names, ms, 9)
        return {names[i].decode(): float(ms[i]) for i in range(n)}
