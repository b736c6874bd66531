"""Non-Limber w(theta) (like.adopt_limber_gg = 0: C_cl_tomo below ell 150) on
the GPU vs the FP64 oracle: the C_gg the w sum reads (non-Limber below each
bin's LMAX, Limber above), LMAX itself, the real-space data vector in both
Legendre modes and with magnification, batches with graph replay, the
reference-named w_gg_tomo(limber = 0), and the refusals (the VJP is in
test_gpu_vjp.py).

Bar: 1e-6 relative per entry with a floor of 1e-3 of the row's largest
magnitude (the FP64 parity metric); LMAX identical.
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


def _problem(seed=3, bmag=None):
    cos = synth.random_draws(1, seed=seed, ragged=False)[0][0]
    p = synth.nonlimber_problem(cos, bmag=bmag)
    rng = np.random.default_rng(seed)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.lens_stretch = 1.0 + rng.normal(0, 0.02, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    return p


@pytest.fixture(scope="module")
def nl():
    p = _problem()
    o = O.Oracle().load(p)
    cl, lmax = o.cl_nonlimber()
    return p, o, cl, lmax, o.datavector()


def test_nonlimber_cl_and_lmax(nl):
    p, o, cl, lmax, dv = nl
    c = Cosmolike().load(p)
    gcl, glmax = c.cl_nonlimber()
    print("lmax", lmax, glmax, "err", rel_err(gcl[:, 1:], cl[:, 1:]))
    assert np.array_equal(glmax, lmax)
    assert rel_err(gcl[:, 1:], cl[:, 1:]) < RTOL


@pytest.mark.parametrize("mode", [LEGENDRE_FOLDED, LEGENDRE_FULL])
def test_nonlimber_datavector(nl, mode):
    p, o, cl, lmax, dv = nl
    c = Cosmolike().load(p, mode)
    got = c.compute_3x2pt_real()
    e = dv_err(got, dv, p.ns, p.nl, p.ntheta)
    print(mode, e)
    assert max(e.values()) < RTOL
    # the w block really differs from the Limber one
    c.set_adopt_limber_gg(1)
    lim = c.compute_3x2pt_real()
    nw = p.nl * p.ntheta
    assert np.array_equal(lim[:-nw], got[:-nw]) or dv_err(lim, got, p.ns, p.nl, p.ntheta)["xi+"] < 1e-14
    assert rel_err(lim[-nw:].reshape(p.nl, -1), got[-nw:].reshape(p.nl, -1)) > 1e-3


def test_nonlimber_with_magnification():
    p = _problem(seed=8, bmag=np.linspace(-0.8, 1.5, 5))
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


def _forward(c, n, B):
    import torch
    out = torch.full((B * n,), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(B, n)


def test_nonlimber_batch_and_graph():
    draws = [_problem(seed=s) for s in (11, 12, 13)]
    draws[1].bmag = np.linspace(0.3, 1.1, 5)
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


def test_nonlimber_mixed_precision(nl):
    p, o, cl, lmax, dv = nl
    c = Cosmolike().load(p)
    c.set_precision("fp32")
    c.batch_upload([p])
    e = dv_err(_forward(c, c.datavector_size(), 1)[0], dv, p.ns, p.nl, p.ntheta)
    print(e)
    assert max(e.values()) < 1e-4


def test_reference_w_gg_tomo_limber0(nl):
    p, o, cl, lmax, dv = nl
    c = Cosmolike.default()
    c.load(p)
    c.set_adopt_limber_gg(1)
    L = c.L
    nt = p.ntheta
    w = dv[-p.nl * nt:].reshape(p.nl, nt)  # w has no calibration
    got = np.array([[L.w_gg_tomo(t, i, i, 0) for t in range(nt)] for i in range(p.nl)])
    assert rel_err(got, w) < RTOL
    c.set_adopt_limber_gg(1)


def test_nonlimber_refusals(nl):
    import torch
    p, *_ = nl
    c = Cosmolike().load(p)
    c.batch_upload([p])
    n = c.datavector_size()
    _forward(c, n, 1)
    bar = torch.ones(n, dtype=torch.float64, device="cuda:0")
    g = torch.zeros(p.nl, dtype=torch.float64, device="cuda:0")
    # the VJP (test_gpu_vjp.py) takes the lens-limit term of the views'
    # rows from the node tangents only
    c.set_vjp_lens_limits(1)
    try:
        with pytest.raises(CosmolikeError, match="non-Limber"):
            c.batch_vjp(bar.data_ptr(), g_lens_dz=g.data_ptr())
    finally:
        c.set_vjp_lens_limits(0)
    q = _problem()
    q.lnPL = None  # no linear P(k,z)
    c2 = Cosmolike().load(q)
    with pytest.raises(CosmolikeError, match="linear P"):
        c2.compute_3x2pt_real()
