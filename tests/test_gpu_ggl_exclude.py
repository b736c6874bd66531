"""gs pair exclusion (init_ggl_exclude, generic_interface.cpp:874-896) on the
GPU: every C_gs column is still integrated, the data vector holds the
included pairs in N_ggl order (redshift_spline.c:199-223).  Against the
oracle at 1e-6 (metric of the FP64 parity tests): single draws in both
Legendre modes and two block orders with the point mass on, FP32 mixed
precision at 1e-4, Fourier space, a batch under graph replay, a ragged batch
of two layouts with different exclusions; the reference-named w_gammat_tomo
and C_gs_tomo_limber return 0 for an excluded pair; the VJP against central
differences of the GPU forward (relative step 1e-4, bar 1e-4 with a floor of
1e-3 of the group's largest difference)."""
import numpy as np
import pytest

from cosmolike_amd import LEGENDRE_FULL, Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu

EXCL = [(0, 0), (0, 3), (2, 9), (4, 1), (4, 9)]


def rel_err(gpu, ref, nt):
    g, r = np.asarray(gpu).reshape(-1, nt), np.asarray(ref).reshape(-1, nt)
    floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(r), floor)
    den[den == 0] = 1.0
    return float((np.abs(g - r) / den).max())


def _prob(seed=3, excl=EXCL, cos=None, ns=None, nl=None, nt=None):
    if ns is None:
        p = synth.full_3x2pt_problem(cos)
    else:
        p = synth.ragged_3x2pt_problem(cos, ns, nl, nt)
    rng = np.random.default_rng(seed)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    p.pm = np.linspace(0.5, 2.0, p.nl)
    p.ggl_exclude = np.array(excl, dtype=np.int32).ravel()
    return p


@pytest.fixture(scope="module")
def prob():
    p = _prob()
    return p, O.Oracle().load(p)


@pytest.mark.parametrize("mode,ord", [(None, (0, 1, 2)), (None, (2, 1, 0)), (LEGENDRE_FULL, (0, 2, 1))])
def test_single_draw(prob, mode, ord):
    p, o = prob
    c = Cosmolike().load(p) if mode is None else Cosmolike().load(p, legendre_mode=mode)
    assert c.datavector_size() == p.ntheta * (2 * 55 + 45 + 5)
    e = rel_err(c.compute_3x2pt_real(ord), o.datavector(ord), p.ntheta)
    print(mode, ord, "err", e)
    assert e < 1e-6
    # gamma_t rows of the included pairs; the reference-named scalars
    gt = c.gammat()
    ref = o.gammat()
    keep = [zl * 10 + zs for zl in range(5) for zs in range(10) if (zl, zs) not in EXCL]
    assert rel_err(gt, ref[keep], p.ntheta) < 1e-6


def test_reference_scalars(prob):
    p, o = prob
    d = Cosmolike.default()
    d.load(p)
    L = d.L
    gt = o.gammat()
    for zl in range(5):
        for zs in range(10):
            v = L.w_gammat_tomo(4, zl, zs, 1)
            c = L.C_gs_tomo_limber(300.0, zl, zs)
            if (zl, zs) in EXCL:
                assert v == 0.0 and c == 0.0
            else:
                assert abs(v - gt[zl * 10 + zs, 4]) <= 1e-6 * np.abs(gt[zl * 10 + zs]).max()
                assert c != 0.0
    d.init_ggl_exclude(None)


def test_mixed_precision_and_fourier(prob):
    p, o = prob
    c = Cosmolike().load(p)
    c.set_precision(1)
    e = rel_err(c.compute_3x2pt_real(), o.datavector(), p.ntheta)
    print("fp32 err", e)
    assert e < 1e-4
    c.set_precision(0)
    c.init_binning_fourier(6, 30, 3000, 1000)
    ells = 30 * (100.0) ** (np.arange(7) / 6)
    ells = np.sqrt(ells[:-1] * ells[1:])
    ref = o.datavector_fourier(ells, 1000, (0, 2, 1))
    got = c.compute_3x2pt_fourier((0, 2, 1))
    assert got.size == ref.size == 6 * (55 + 45 + 5)
    assert rel_err(got, ref, 6) < 1e-6


def test_batches(prob):
    import torch
    p, o = prob
    spec = synth.random_draws(3, seed=11, ragged=False)
    draws = [_prob(20 + i, cos=s[0]) for i, s in enumerate(spec)]
    c = Cosmolike().load(draws[0])
    c.batch_upload(draws)
    n = c.datavector_size()
    out = torch.full((3 * n,), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    eager = out.cpu().numpy()
    for i, d in enumerate(draws):
        assert rel_err(eager[i * n:(i + 1) * n], O.Oracle().load(d).datavector(), d.ntheta) < 1e-6
    c.batch_use_graph(True)
    try:
        for _ in range(2):
            out.fill_(float("nan"))
            c.batch_forward(out.data_ptr())
            c.synchronize()
            assert np.array_equal(out.cpu().numpy(), eager)
    finally:
        c.batch_use_graph(False)
    # ragged: two layouts with different exclusions
    cos = synth.random_draws(2, seed=12, ragged=False)
    probs = [_prob(30, cos=cos[0][0], ns=4, nl=3, nt=20, excl=[(1, 2), (2, 0)]),
             _prob(31, cos=cos[1][0], ns=6, nl=2, nt=24, excl=[(0, 5)])]
    lays = []
    for q in probs:
        c.load(q)
        lays.append(c.capture_layout())
    c.batch_upload_layouts(probs, lays)
    offs = c.batch_offsets()
    out = torch.full((offs[-1],), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    got = out.cpu().numpy()
    for i, q in enumerate(probs):
        ref = O.Oracle().load(q).datavector()
        assert offs[i + 1] - offs[i] == ref.size
        assert rel_err(got[offs[i]:offs[i + 1]], ref, q.ntheta) < 1e-6


def test_vjp_vs_fd():
    import torch
    from cosmolike_amd.autograd import projection_3x2pt
    spec = synth.random_draws(2, seed=13, ragged=False)
    draws = [_prob(40 + i, cos=s[0]) for i, s in enumerate(spec)]
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP.ravel() for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(5).normal(size=(2, c.datavector_size())))
    fwd = lambda Q: projection_3x2pt(c, Q["lnP"], Q["src_dz"], Q["lens_dz"], Q["b1"], Q["ia_A"], Q["shear_m"])
    leaves = {k: x.clone().requires_grad_(True) for k, x in P.items()}
    (fwd(leaves) * v).sum().backward()
    torch.cuda.synchronize()
    for group in ("src_dz", "b1", "shear_m"):
        g = leaves[group].grad.detach().cpu().numpy()
        fd = np.zeros_like(g)
        for j in range(g.shape[1]):
            h = 1e-4 * np.abs(P[group][:, j].cpu().numpy())
            ht = torch.tensor(h, dtype=torch.float64, device=dev)
            xp = P[group].clone(); xp[:, j] += ht
            xm = P[group].clone(); xm[:, j] -= ht
            with torch.no_grad():
                fd[:, j] = ((fwd(dict(P, **{group: xp})) - fwd(dict(P, **{group: xm}))) * v).sum(1).cpu().numpy() / (2 * h)
        err = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
        print(group, "max rel err", err.max())
        assert err.max() < 1e-4, group
