"""One-loop galaxy bias (b2, bs2, b3, bK with the FAST-PT bias tables of
set_bias_PS; cosmolike/cosmo2D.c:1970-2057, 2560-2603) on the GPU vs the FP64
oracle: Limber tables, the real-space data vector in both Legendre modes,
batches with graph replay, FP32 mode, with the non-Limber w, and the
refusal together with TATT (the VJP is in test_gpu_vjp.py).

Bar: 1e-6 relative per entry with a floor of 1e-3 of the row's largest
magnitude (the FP64 parity metric); FP32 mode 1e-4.
"""
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


def _problem(seed=4):
    cos = synth.random_draws(1, seed=seed, ragged=False)[0][0]
    p = synth.oneloop_problem(cos)
    rng = np.random.default_rng(seed)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    return p


@pytest.fixture(scope="module")
def ol():
    p = _problem()
    o = O.Oracle().load(p)
    return p, o, o.datavector(), o.cl_tables()


def test_oneloop_tables(ol):
    p, o, dv, (ss, gs, gg) = ol
    c = Cosmolike().load(p)
    g_ss, g_gs, g_gg = c.cl_tables(p.n_ell)
    e = {"ss": rel_err(g_ss, ss), "gs": rel_err(g_gs, gs), "gg": rel_err(g_gg, gg)}
    print(e)
    assert max(e.values()) < RTOL
    ells = np.arange(1.0, 20.0)
    for probe in ("gs", "gg"):
        assert rel_err(c.cl_nointerp(probe, ells), o.cl_nointerp(probe, ells)) < RTOL, probe


@pytest.mark.parametrize("mode", [LEGENDRE_FOLDED, LEGENDRE_FULL])
def test_oneloop_datavector(ol, mode):
    p, o, dv, _ = ol
    c = Cosmolike().load(p, mode)
    e = dv_err(c.compute_3x2pt_real(), dv, p.ns, p.nl, p.ntheta)
    print(mode, e)
    assert max(e.values()) < RTOL


def _forward(c, n, B):
    import torch
    out = torch.full((B * n,), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(B, n)


def test_oneloop_batch_and_graph():
    draws = [_problem(seed=s) for s in (21, 22, 23)]
    draws[2].b2 = np.zeros(draws[2].nl)  # one draw without the one-loop terms
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


def test_oneloop_fp32(ol):
    p, o, dv, _ = ol
    c = Cosmolike().load(p)
    c.set_precision("fp32")
    c.batch_upload([p])
    e = dv_err(_forward(c, c.datavector_size(), 1)[0], dv, p.ns, p.nl, p.ntheta)
    print(e)
    assert max(e.values()) < 1e-4


def test_oneloop_with_nonlimber_w():
    p = _problem(seed=6)
    q = synth.nonlimber_problem(synth.random_draws(1, seed=6, ragged=False)[0][0])
    p.plin_log10k, p.plin_z, p.lnPL, p.limber_gg = q.plin_log10k, q.plin_z, q.lnPL, 0
    o = O.Oracle().load(p)
    ref = o.datavector()
    cl, lmax = o.cl_nonlimber()
    c = Cosmolike().load(p)
    gcl, glmax = c.cl_nonlimber()
    assert np.array_equal(glmax, lmax)
    assert rel_err(gcl[:, 1:], cl[:, 1:]) < RTOL
    e = dv_err(c.compute_3x2pt_real(), ref, p.ns, p.nl, p.ntheta)
    print(e)
    assert max(e.values()) < RTOL


def test_oneloop_refusals(ol):
    """TATT together with the one-loop bias is refused (the TATT gs one-loop
    term multiplies W_K alone); the VJP of one-loop batches is tested
    against finite differences in test_gpu_vjp.py."""
    p, *_ = ol
    t = synth.tatt_problem()
    t.b2, t.b3, t.bk = p.b2, p.b3, p.bk
    t.bias_ps, t.bias_ps_kmin, t.bias_ps_kmax, t.bias_ps_sigma4 = p.bias_ps, p.bias_ps_kmin, p.bias_ps_kmax, \
        p.bias_ps_sigma4
    c2 = Cosmolike().load(t)
    with pytest.raises(CosmolikeError, match="TATT"):
        c2.compute_3x2pt_real()
