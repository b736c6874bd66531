"""Likelihood end on the GPU vs the FP64 oracle: point mass on gamma_t, the
mask, chi^2 with the inverse covariance and its gradient, and the chain of
that gradient through the adjoint (d chi^2 / d parameters vs finite
differences of chi^2).

Bars: data vectors 1e-6 relative (metric of the FP64 parity tests); chi^2
1e-9 relative; gradients of chi^2 1e-4 against central differences
(relative step 1e-4, 1e-5 absolute where the parameter is 0).
"""
import numpy as np
import pytest

from cosmolike_amd import Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu


def rel_err(gpu, ref, nt):
    g = gpu.reshape(-1, nt)
    r = ref.reshape(-1, nt)
    floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(r), floor)
    den[den == 0] = 1.0
    return float((np.abs(g - r) / den).max())


@pytest.fixture(scope="module")
def like():
    import torch
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(3, seed=31, ragged=False)):
        p = synth.full_3x2pt_problem(cos)
        rng = np.random.default_rng(50 + i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.pm = np.abs(rng.normal(0, 2.0, p.nl))
        draws.append(p)
    n = 26 * (2 * 55 + 50 + 5)
    rng = np.random.default_rng(7)
    mask = (rng.random(n) > 0.2).astype(np.int32)
    ref0 = O.Oracle().load(draws[0]).datavector()
    data = ref0 * (1 + 0.05 * rng.normal(size=n))
    sig = 0.1 * np.abs(ref0) + 1e-3 * np.abs(ref0).max()
    A = rng.normal(size=(n, 40))
    cov = np.diag(sig ** 2) + (A * sig[:, None]) @ (A * sig[:, None]).T / 400
    keep = mask == 1
    icov = np.zeros((n, n))
    icov[np.ix_(keep, keep)] = np.linalg.inv(cov[np.ix_(keep, keep)])
    c = Cosmolike().load(draws[0])
    c.like_set_mask(mask)
    c.like_set_data(data)
    c.like_set_inv_cov(icov)
    c.batch_upload(draws)
    refs = []
    for d in draws:
        o = O.Oracle().load(d)
        o.set_mask(mask)
        refs.append(o.datavector())
    return c, draws, mask, data, icov, refs, n, torch


def test_masked_point_mass_datavector(like):
    c, draws, mask, data, icov, refs, n, torch = like
    out = torch.empty(len(draws) * n, dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    got = out.view(len(draws), n).cpu().numpy()
    for i in range(len(draws)):
        assert np.all(got[i][mask == 0] == 0)
        e = rel_err(got[i], refs[i], 26)
        print("draw", i, "err", e)
        assert e < 1e-6
    # single-draw drop-in path
    c1 = Cosmolike().load(draws[1])
    c1.like_set_mask(mask)
    assert rel_err(c1.compute_3x2pt_real(), refs[1], 26) < 1e-6


def test_chi2_batch_and_single(like):
    c, draws, mask, data, icov, refs, n, torch = like
    B = len(draws)
    out = torch.empty(B * n, dtype=torch.float64, device="cuda:0")
    chi2 = torch.empty(B, dtype=torch.float64, device="cuda:0")
    grad = torch.empty(B * n, dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.batch_chi2(out.data_ptr(), chi2.data_ptr(), grad.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    dv = out.view(B, n).cpu().numpy()
    for i in range(B):
        ref = O.chi2(dv[i], data, icov, mask)
        print("chi2", i, chi2[i].item(), ref)
        assert abs(chi2[i].item() - ref) <= 1e-9 * ref
        r = np.where(mask == 1, dv[i] - data, 0.0)
        g_ref = 2.0 * mask * (icov @ r)
        g = grad.view(B, n)[i].cpu().numpy()
        assert np.abs(g - g_ref).max() <= 1e-9 * np.abs(g_ref).max()
    assert abs(c.chi2(dv[0]) - O.chi2(dv[0], data, icov, mask)) <= 1e-9 * O.chi2(dv[0], data, icov, mask)


def test_chi2_gradient_chain_vs_fd(like):
    """d chi^2 / d (src_dz, shear_m, b1): chi^2 gradient as the VJP cotangent."""
    c, draws, mask, data, icov, refs, n, torch = like
    B = len(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    keys = ("lnP", "src_dz", "lens_dz", "b1", "ia_A", "shear_m")
    out = torch.empty(B * n, dtype=torch.float64, device=dev)
    chi2 = torch.empty(B, dtype=torch.float64, device=dev)
    grad = torch.empty(B * n, dtype=torch.float64, device=dev)

    def run(PP):
        c.batch_set_params(*[PP[k].data_ptr() for k in keys])
        c.batch_forward(out.data_ptr())
        c.batch_chi2(out.data_ptr(), chi2.data_ptr(), grad.data_ptr())
        c.synchronize()
        torch.cuda.synchronize()
        return chi2.cpu().numpy().copy()

    run(P)
    g = {k: torch.empty_like(x) for k, x in P.items()}
    c.batch_vjp(grad.data_ptr(), *[g[k].data_ptr() for k in keys])
    c.synchronize()
    torch.cuda.synchronize()
    for name, j in (("src_dz", 2), ("src_dz", 9), ("shear_m", 4), ("b1", 3)):
        x = P[name]
        h = 1e-4 * x[:, j].abs()
        h = torch.where(h > 0, h, torch.full_like(h, 1e-5))
        xp, xm = x.clone(), x.clone()
        xp[:, j] += h
        xm[:, j] -= h
        fd = (run(dict(P, **{name: xp})) - run(dict(P, **{name: xm}))) / (2 * h.cpu().numpy())
        got = g[name][:, j].cpu().numpy()
        print(name, j, got, fd)
        assert np.all(np.abs(got - fd) <= 1e-4 * np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max()))
    run(P)
