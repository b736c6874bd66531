"""TATT intrinsic alignments (IA_MODEL = TATT) on the GPU vs the FP64 oracle:
E- and B-mode C_ss tables, the TATT C_gs, xi+- from EE +- BB, gamma_t and w
of the real-space 3x2pt data vector, direct-GL C_ell, batches, the precision
modes (TATT ss / gs tiles stay FP64) and the VJP's refusal above 10 source
bins (the VJP itself is in test_gpu_vjp.py).

The FAST-PT IA spectra are synthetic tables of the set_IA_PS shape (no
FAST-PT here); the oracle reads the same tables.

Bar: 1e-6 relative per entry with a floor of 1e-3 of the (probe, pair) row's
largest magnitude (the FP64 parity metric); precision modes 1e-4 (FP32 and
BF16X3), their bars on NLA draws.
"""
import copy

import numpy as np
import pytest

from cosmolike_amd import Cosmolike, CosmolikeError, LEGENDRE_FOLDED, LEGENDRE_FULL, synth

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


def dv_err(dv, ref, ns, nl, nt):
    n_ss, n_gs = ns * (ns + 1) // 2, ns * nl
    worst, o = {}, 0
    for name, rows in (("xi+", n_ss), ("xi-", n_ss), ("gammat", n_gs), ("w", nl)):
        worst[name] = rel_err(dv[o:o + rows * nt].reshape(rows, nt), ref[o:o + rows * nt].reshape(rows, nt))
        o += rows * nt
    return worst


def _problem(seed=5, **kw):
    cos = synth.random_draws(1, seed=seed, ragged=False)[0][0]
    p = synth.tatt_problem(cos, **kw)
    rng = np.random.default_rng(seed)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    return p


@pytest.fixture(scope="module")
def tatt():
    p = _problem()
    o = O.Oracle().load(p)
    return p, o, o.datavector(), o.cl_tables(), o.cl_tables_bb()


def test_tatt_cl_tables(tatt):
    p, o, dv, (ss, gs, gg), bb = tatt
    c = Cosmolike().load(p)
    g_ss, g_gs, g_gg = c.cl_tables(p.n_ell)
    g_bb = c.cl_tables_bb(p.n_ell)
    e = {"ss": rel_err(g_ss, ss), "bb": rel_err(g_bb, bb), "gs": rel_err(g_gs, gs), "gg": rel_err(g_gg, gg)}
    print("tables", e)
    assert np.abs(bb).max() > 0  # the B mode is really exercised
    assert max(e.values()) < RTOL


@pytest.mark.parametrize("mode", [LEGENDRE_FOLDED, LEGENDRE_FULL])
def test_tatt_datavector(tatt, mode):
    p, o, dv, *_ = tatt
    c = Cosmolike().load(p, mode)
    got = c.compute_3x2pt_real()
    e = dv_err(got, dv, p.ns, p.nl, p.ntheta)
    print("mode", mode, e)
    assert max(e.values()) < RTOL
    # xi+ and xi- differ from their NLA values by more than the bar: the
    # B mode and the TATT terms enter them
    q = copy.copy(p)
    q.ia_model = 0
    nla = O.Oracle().load(q).datavector()
    m = p.ns * (p.ns + 1) // 2 * p.ntheta
    assert rel_err(dv[:m], nla[:m]) > 1e-3


def test_tatt_nointerp(tatt):
    p, o, *_ = tatt
    c = Cosmolike().load(p)
    ells = np.exp(np.linspace(np.log(3.0), np.log(5000.0), 25))
    for probe in ("ss", "gs", "gg"):
        assert rel_err(c.cl_nointerp(probe, ells), o.cl_nointerp(probe, ells)) < RTOL, probe
    assert rel_err(c.cl_nointerp("ss_bb", ells), o.cl_ss_bb_nointerp(ells)) < RTOL


def test_tatt_per_bin_amplitudes():
    # ia_redshift_evol 2: A1, A2 and b_TA given per source bin
    p = _problem(seed=9)
    p.ia_mode = 2
    rng = np.random.default_rng(9)
    a1, a2, bta = rng.normal(0.5, 0.3, p.ns), rng.normal(-0.5, 0.4, p.ns), rng.normal(0.8, 0.3, p.ns)
    o = O.Oracle().load(p)
    o.set_ia(2, A_bin=a1)
    o.set_ia_tatt(1, A2_bin=a2, bta_bin=bta)
    ref = o.datavector()
    c = Cosmolike().load(p)
    c.set_nuisance_IA(a1, a2, bta)
    got = c.compute_3x2pt_real()
    e = dv_err(got, ref, p.ns, p.nl, p.ntheta)
    print(e)
    assert max(e.values()) < RTOL


def _forward(c, n, B):
    import torch
    out = torch.full((B * n,), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(B, n)


def test_tatt_batch_and_graph():
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(3, seed=13, ragged=False)):
        q = synth.tatt_problem(cos, A2=-0.8 + 0.5 * i, eta2=0.3 * i, bta=0.4 + 0.3 * i)
        q.shear_m = np.full(q.ns, 0.01 * i)
        draws.append(q)
    c = Cosmolike().load(draws[0])
    c.batch_upload(draws)
    n = c.datavector_size()
    got = _forward(c, n, len(draws))
    for i, q in enumerate(draws):
        e = dv_err(got[i], O.Oracle().load(q).datavector(), q.ns, q.nl, q.ntheta)
        print(i, e)
        assert max(e.values()) < RTOL, i
    c.batch_use_graph(True)
    for _ in range(2):
        assert np.array_equal(_forward(c, n, len(draws)), got)
    c.batch_use_graph(False)


@pytest.mark.parametrize("mode,tol", [("fp32", 1e-4), ("bf16x3", 1e-4)])
def test_tatt_precision_modes(tatt, mode, tol):
    p, o, dv, *_ = tatt
    c = Cosmolike().load(p)
    c.set_precision(mode)
    c.batch_upload([p])
    got = _forward(c, c.datavector_size(), 1)[0]
    e = dv_err(got, dv, p.ns, p.nl, p.ntheta)
    print(mode, e)
    assert max(e.values()) < tol


def test_tatt_vjp_refused_above_10_source_bins():
    import torch
    cos = synth.random_draws(1, seed=5, ragged=False)[0][0]
    p = synth.ragged_3x2pt_problem(cos, 11, 2, 20)
    q = synth.tatt_problem(cos)
    p.ia_model, p.ia2_A, p.ia2_eta, p.ia_bta = q.ia_model, q.ia2_A, q.ia2_eta, q.ia_bta
    p.ia_ps, p.ia_ps_kmin, p.ia_ps_kmax = q.ia_ps, q.ia_ps_kmin, q.ia_ps_kmax
    c = Cosmolike().load(p)
    c.batch_upload([p])
    n = c.datavector_size()
    _forward(c, n, 1)
    bar = torch.ones(n, dtype=torch.float64, device="cuda:0")
    g = torch.zeros(p.ns, dtype=torch.float64, device="cuda:0")
    with pytest.raises(CosmolikeError, match="TATT"):
        c.batch_vjp(bar.data_ptr(), g_src_dz=g.data_ptr())


def test_tatt_needs_tables():
    p = _problem()
    p.ia_ps = None
    c = Cosmolike()
    c.init_ntable(p.n_ell, 225, p.lmax, 0)
    c.init_binning_real_space(p.ntheta, p.theta_min_arcmin, p.theta_max_arcmin)
    c.init_IA(1, 3)
    c.set_source_sample(p.source)
    c.set_lens_sample(p.lens)
    p.ia_model = 0  # set_draw without the tables
    c.set_draw(p)
    with pytest.raises(CosmolikeError, match="FAST-PT"):
        c.compute_3x2pt_real()
