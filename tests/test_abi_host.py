"""C-ABI surface and host-side logic of libcosmolike_hip.so, without a GPU.

The library's setters and validation run on the host; a context touches the
device only when something is computed, so these run in the CPU suite.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from cosmolike_amd import Cosmolike, CosmolikeError, _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "cosmolike_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted({n for n in names if n not in ("if", "while", "return", "sizeof")})


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_names_present():
    names = set(header_functions())
    for n in ("xi_pm_tomo", "w_gammat_tomo", "w_gg_tomo", "C_ss_tomo_limber", "C_gs_tomo_limber",
              "C_gg_tomo_limber", "C_ss_tomo_limber_nointerp", "C_gs_tomo_limber_nointerp",
              "C_gg_tomo_limber_nointerp"):
        assert n in names


@pytest.fixture()
def ctx():
    c = Cosmolike(device=0)
    yield c
    c.close()


def test_setters_validate(ctx):
    p = synth.shear_only_problem()
    ctx.load(p)
    assert ctx.datavector_size() == 2 * 15 * 20
    with pytest.raises(CosmolikeError) as e:
        ctx.init_binning_real_space(40, 2.5, 250.0)  # more than 32 theta bins
    assert e.value.code == -1
    with pytest.raises(CosmolikeError) as e:
        ctx.init_ntable(300, 1, 50000, 1)             # N_a < 2
    assert e.value.code == -1
    ctx.init_ntable(300, 225, 50000, 3)               # high_def_integration != 0 is accepted
    ctx.init_ntable(p.n_ell, 225, p.lmax, 0)
    bad = p.chi.copy()
    bad[3] = np.nan
    with pytest.raises(CosmolikeError) as e:
        ctx.set_distances(p.chi_z, bad)
    assert e.value.code == -2
    with pytest.raises(CosmolikeError) as e:
        ctx.set_nuisance_shear_calib(np.zeros(4))     # 5 source bins
    assert e.value.code == -1
    with pytest.raises(CosmolikeError) as e:
        ctx.set_nuisance_linear_bias(np.ones(5))      # no lens sample
    assert e.value.code == -3
    tab = p.source.copy()
    tab[10, 0] = tab[9, 0]                            # z column not monotonic
    with pytest.raises(CosmolikeError) as e:
        ctx.set_source_sample(tab)
    assert e.value.code == -4
    tab = p.source.copy()
    tab[:, 2] = 0.0                                   # empty bin
    with pytest.raises(CosmolikeError) as e:
        ctx.set_source_sample(tab)
    assert e.value.code == -4
    with pytest.raises(CosmolikeError) as e:
        ctx.set_precision(3)                          # no such precision mode
    assert e.value.code == -1
    for m in ("fp32", "bf16x3", "fp64"):
        ctx.set_precision(m)
    for bad in ([0, 16, 16, 16], [16, 65, 16, 16]):
        with pytest.raises(CosmolikeError) as e:
            ctx.set_tile_rows(bad)                    # rows per tile outside 1..64
        assert e.value.code == -1
    ctx.set_tile_rows([32, 32, 16, 64])
    assert ctx.tile_rows() == [32, 32, 16, 64]


def test_tatt_setters_validate(ctx):
    p = synth.tatt_problem()
    with pytest.raises(CosmolikeError) as e:
        ctx.init_IA(2, 3)                             # no such IA model
    assert e.value.code == -1
    ctx.load(p)                                       # TATT layout and tables
    with pytest.raises(CosmolikeError) as e:
        ctx.set_IA_PS(p.ia_ps, 1.0, 0.5)              # kmax < kmin
    assert e.value.code == -1
    bad = p.ia_ps.copy().reshape(12, -1)
    bad[3, 7] = np.nan
    with pytest.raises(CosmolikeError) as e:
        ctx.set_IA_PS(bad, p.ia_ps_kmin, p.ia_ps_kmax)
    assert e.value.code == -2
    with pytest.raises(CosmolikeError) as e:
        ctx.set_nuisance_IA([0.5, 0.0], [np.nan, 0.0], [1.0])
    assert e.value.code == -2
    ctx.init_IA(1, 2)
    with pytest.raises(CosmolikeError) as e:
        ctx.set_nuisance_IA(np.ones(3), np.ones(3), np.ones(3))  # fewer than p.ns
    assert e.value.code == -1
    ctx.set_nuisance_IA(np.ones(p.ns), np.ones(p.ns), np.ones(p.ns))


def test_new_setters_validate(ctx):
    p = synth.full_3x2pt_problem()
    ctx.load(p)
    with pytest.raises(CosmolikeError) as e:
        ctx.init_probes("5x2pt")                      # only 3x2pt and 6x2pt
    assert e.value.code == -1
    with pytest.raises(CosmolikeError) as e:
        ctx.set_adopt_limber_gg(2)
    assert e.value.code == -1
    bad = p.lnP.copy()
    bad[5] = np.nan
    with pytest.raises(CosmolikeError) as e:
        ctx.set_linear_power_spectrum(p.log10k, p.pk_z, bad)
    assert e.value.code == -2
    with pytest.raises(CosmolikeError) as e:
        ctx.set_nuisance_nonlinear_bias(p.b1, np.zeros(3))  # 5 lens bins
    assert e.value.code == -1
    with pytest.raises(CosmolikeError) as e:
        ctx.set_nuisance_nonlocal_bias(np.zeros(5), np.full(5, np.nan))
    assert e.value.code == -2
    tab, kmin, kmax, s4 = synth.fpt_bias_tables(synth.Cosmology(), n=64)
    with pytest.raises(CosmolikeError) as e:
        ctx.set_bias_PS(tab, kmax, kmin, s4)            # kmax < kmin
    assert e.value.code == -1
    with pytest.raises(CosmolikeError) as e:
        ctx.init_cmb_cross_correlation(8, 4000, float("nan"))
    assert e.value.code == -1
    # 6x2pt adds (nl + ns) ntheta + N_ell(Fourier) entries to the real-space vector
    n3 = ctx.datavector_size()
    ctx.init_binning_fourier(20, 30, 3000, 3000)
    ctx.init_probes("6x2pt")
    assert ctx.datavector_size() == n3 + (p.nl + p.ns) * p.ntheta + 20
    ctx.init_probes("3x2pt")
    assert ctx.datavector_size() == n3


def test_kk_bandpower_setter(ctx):
    """init_cmb_auto_bandpower: validation as IPCMB::set_kk_binning_bandpower
    (nbins, lmin, lmax > 0), NaN refusal, and the kk block size becomes nbins."""
    p = synth.sixpt_problem()
    ctx.load(p)
    n6 = ctx.datavector_size()
    assert ctx.kk_bandpower_info() == (0, 0, 0, 1.0)
    M = np.ones((4, 100 - 8 + 1))
    for bad in ((0, 8, 100), (4, 0, 100), (4, 50, 10)):
        with pytest.raises((CosmolikeError, ValueError)):
            ctx.init_cmb_auto_bandpower(*bad, np.ones(max(bad[0], 1) * max(bad[2] - bad[1] + 1, 1)))
    with pytest.raises(CosmolikeError) as e:
        ctx.init_cmb_auto_bandpower(4, 8, 100, M, alpha=0.0)
    assert e.value.code == -1
    Mn = M.copy()
    Mn[2, 7] = np.nan
    with pytest.raises(CosmolikeError) as e:
        ctx.init_cmb_auto_bandpower(4, 8, 100, Mn)
    assert e.value.code == -2
    with pytest.raises(ValueError):
        ctx.init_cmb_auto_bandpower(4, 8, 100, M, np.zeros(3))
    assert ctx.kk_bandpower_info()[0] == 0 and ctx.datavector_size() == n6
    ctx.init_cmb_auto_bandpower(4, 8, 100, M, np.zeros(4), alpha=0.9)
    assert ctx.kk_bandpower_info() == (4, 8, 100, 0.9)
    assert ctx.datavector_size() == n6 - p.kk_nells + 4


def test_datavector_sizes(ctx):
    p = synth.full_3x2pt_problem()
    ctx.load(p)
    assert ctx.datavector_size() == 26 * (2 * 55 + 50 + 5) == 4290


def test_fourier_binning_validates(ctx):
    p = synth.full_3x2pt_problem()
    ctx.load(p)
    assert ctx.datavector_size_fourier() == 0
    ctx.init_binning_fourier(15, 30, 3000, 1000)
    assert ctx.datavector_size_fourier() == 15 * (55 + 50 + 5)
    for bad in ((0, 30, 3000, 1000), (15, 0, 3000, 1000), (15, 3000, 30, 1000)):
        with pytest.raises(CosmolikeError) as e:
            ctx.init_binning_fourier(*bad)
        assert e.value.code == -1
    assert ctx.datavector_size_fourier() == 15 * 110
    with pytest.raises(ValueError):
        ctx.compute_3x2pt_fourier(mask=np.ones(5))


@pytest.mark.parametrize("ord", [(1, 2, 0), (2, 0, 1), (0, 0, 1), (0, 1, 3), (-1, 0, 1)])
def test_orders_that_do_not_tile_are_refused(ctx, ord):
    """The reference's block-start formula (generic_interface.hpp:781-787)
    tiles the data vector only for involutive orders; any other order is
    refused before any device work, for real and Fourier space alike."""
    p = synth.full_3x2pt_problem()
    ctx.load(p)
    ctx.init_binning_fourier(3, 30, 3000, 1000)
    for fn in (lambda: ctx.compute_3x2pt_real(ord), lambda: ctx.compute_3x2pt_fourier(ord),
               lambda: ctx.like_set_mask(np.ones(ctx.datavector_size(), np.int32), ord)):
        with pytest.raises(CosmolikeError) as e:
            fn()
        assert e.value.code == -1


def test_compute_without_device_fails_cleanly(ctx):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    p = synth.shear_only_problem()
    ctx.load(p)
    out = np.full(ctx.datavector_size(), 7.0)
    o = (ctypes.c_int * 3)(0, 1, 2)
    rc = ctx.L.clh_compute_3x2pt_real(ctx.h, o, out.ctypes.data_as(_lib.D), len(out))
    assert rc == -5
    assert np.all(out == 7.0)


def test_layout_capture_is_a_snapshot(ctx):
    p = synth.full_3x2pt_problem()
    ctx.load(p)
    a = ctx.capture_layout()
    assert a.datavector_size() == 4290
    ctx.init_binning_real_space(20, 2.5, 250.0)
    b = ctx.capture_layout()
    assert b.datavector_size() == 20 * (2 * 55 + 50 + 5)
    assert a.datavector_size() == 4290  # unchanged by later setters
    del a, b


EXCL = [(0, 0), (0, 3), (2, 9), (4, 1), (4, 9)]


def test_ggl_exclude_sizes_and_maps(ctx):
    """init_ggl_exclude (generic_interface.cpp:874-896): the gamma_t block
    loses the excluded pairs (tomo.ggl_Npowerspectra, :775-802); the pair maps
    test_zoverlap / N_ggl / ZL / ZS (redshift_spline.c:115-223) follow."""
    p = synth.full_3x2pt_problem()
    ctx.load(p)
    nt = p.ntheta
    assert ctx.datavector_size() == nt * (2 * 55 + 50 + 5)
    ctx.init_binning_fourier(4, 30, 3000, 1000)
    ctx.init_ggl_exclude(EXCL)
    assert ctx.n_ggl == 45
    assert ctx.datavector_size() == nt * (2 * 55 + 45 + 5)
    assert ctx.datavector_size_fourier() == 4 * (55 + 45 + 5)
    for zl in range(5):
        for zs in range(10):
            assert ctx.test_zoverlap(zl, zs) == (0 if (zl, zs) in EXCL else 1)
    with pytest.raises(CosmolikeError):
        ctx.test_zoverlap(5, 0)
    # a mask of the old size is dropped; the flat form is accepted too
    ctx.init_ggl_exclude(np.array(EXCL[:2]).ravel())
    assert ctx.datavector_size() == nt * (2 * 55 + 48 + 5)
    ctx.init_ggl_exclude(None)
    assert ctx.datavector_size() == nt * (2 * 55 + 50 + 5)
    # the reference-named maps on the default context
    d = Cosmolike.default()
    d.load(p)
    d.init_ggl_exclude(EXCL)
    L = d.L
    n = 0
    for zl in range(5):
        for zs in range(10):
            inc = (zl, zs) not in EXCL
            assert L.test_zoverlap(zl, zs) == int(inc)
            assert L.N_ggl(zl, zs) == (n if inc else -1)
            if inc:
                assert (L.ZL(n), L.ZS(n)) == (zl, zs)
                n += 1
    assert n == 45 and L.ZL(45) == -1 and L.N_ggl(0, 10) == -1 and L.test_zoverlap(-1, 0) == -1
    d.init_ggl_exclude(None)
