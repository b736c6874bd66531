"""The batched configurations at their stated sizes.

- 512 draws with their own cosmology (Omega_m, sigma8, h uniform) and their
  own survey layout (3-10 source x 2-6 lens bins, 20-30 theta bins) in one
  batch replayed from a captured hipGraph: every draw against the FP64 oracle
  (1e-6), every replay bit-identical to the eager launches.
- The VJP at batch 64 on the 10 + 5 bin 3x2pt layout, draws with their own
  cosmology and nuisance values (b_mag != 0 on a third of them), against an
  N(0,1) cotangent: every scalar group (source and lens shifts, b1, A_IA, m)
  and a random ln P(k,z) direction against central differences of the GPU
  forward (relative step 1e-4, absolute 1e-5 at zero), 1e-4 with a floor of
  1e-3 of the group's largest difference.  The lens shifts also move the lens
  integration limits (2 |dz| margins), and with them the GL nodes; the
  forward then jumps where a node crosses a knot of the G(z) or chi(z) table
  (the growth rate and dchi/dz are their interval slopes), a few times per
  1e-6 of dz.  Their difference is the median of the slopes over eight steps
  of 2e-8 about the value, which leaves out the steps holding a jump.

The draws are built by a pool of fresh (spawned) interpreters.
"""
import multiprocessing as mp

import numpy as np
import pytest

from cosmolike_amd import Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu


def _build(item):
    kind, spec = item
    if kind == "ragged":
        c, ns, nl, nt = spec
        return synth.ragged_3x2pt_problem(c, ns, nl, nt)
    return synth.full_3x2pt_problem(spec)


def _build_all(items):
    with mp.get_context("spawn").Pool(16) as pool:
        return pool.map(_build, items, chunksize=4)


def _dv_err(gpu, ref, ns, nl, nt):
    n_ss, n_gs = ns * (ns + 1) // 2, ns * nl
    worst, o = 0.0, 0
    for rows in (n_ss, n_ss, n_gs, nl):
        g = gpu[o:o + rows * nt].reshape(rows, nt)
        r = ref[o:o + rows * nt].reshape(rows, nt)
        floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
        worst = max(worst, float((np.abs(g - r) / np.maximum(np.abs(r), floor)).max()))
        o += rows * nt
    return worst


def test_ragged_512_draws_graph_vs_oracle():
    import torch
    spec = synth.random_draws(512, seed=20261204, ragged=True)
    probs = _build_all([("ragged", tuple(s)) for s in spec])
    rng = np.random.default_rng(4)
    for p in probs:
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.005, p.nl)
    c = Cosmolike()
    cache, lays = {}, []
    for p in probs:
        key = (p.ns, p.nl, p.ntheta)
        if key not in cache:
            c.load(p)
            cache[key] = c.capture_layout()
        lays.append(cache[key])
    print("draws", len(probs), "layouts", len(cache))
    c.batch_upload_layouts(probs, lays)
    offs = c.batch_offsets()
    n = offs[-1]
    out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    eager = out.cpu().numpy()
    c.batch_use_graph(True)
    try:
        for _ in range(2):
            out.fill_(float("nan"))
            c.batch_forward(out.data_ptr())
            c.synchronize()
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy(), eager)
    finally:
        c.batch_use_graph(False)
    # the oracle keeps one context per layout (its Legendre kernels) and
    # takes each draw's cosmology and nuisance values
    oracles = {}
    worst = 0.0
    for i, p in enumerate(probs):
        key = (p.ns, p.nl, p.ntheta)
        if key not in oracles:
            oracles[key] = O.Oracle().load(p)
        o = oracles[key]
        o.set_cosmology(p.omega_m, p.chi_z, p.chi, p.G_z, p.G, p.log10k, p.pk_z, p.lnP)
        o.set_nuisance(p.src_dz, p.lens_dz, p.lens_stretch, p.b1, p.bmag, p.shear_m)
        o.touch()
        e = _dv_err(eager[offs[i]:offs[i + 1]], o.datavector(), p.ns, p.nl, p.ntheta)
        worst = max(worst, e)
        assert e < 1e-6, (i, p.ns, p.nl, p.ntheta, e)
    print("worst relative error over 512 draws", worst)


@pytest.fixture(scope="module")
def vjp64():
    import torch
    spec = synth.random_draws(64, seed=5, ragged=False)
    draws = _build_all([("full", s[0]) for s in spec])
    rng = np.random.default_rng(64)
    for i, p in enumerate(draws):
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.005, p.nl)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.b1 = np.linspace(1.2, 2.0, p.nl) * (1 + 0.1 * rng.normal())
        p.ia_A = 0.5 + 0.2 * rng.normal()
        p.bmag = rng.normal(0, 1.0, p.nl) if i % 3 == 0 else np.zeros(p.nl)
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP.ravel() for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(42).normal(size=(len(draws), c.datavector_size())))
    return c, draws, P, v


def _fwd(c, P):
    from cosmolike_amd.autograd import projection_3x2pt
    return projection_3x2pt(c, P["lnP"], P["src_dz"], P["lens_dz"], P["b1"], P["ia_A"], P["shear_m"])


def _median_slope(c, P, group, flat, j, step=2e-8, n=8):
    """Median over k of (F(x + (k+1) step) - F(x + k step)) / step, k = -n/2 .. n/2 - 1,
    F = v . forward, per draw."""
    import torch
    v = P["_v"]
    F = []
    for k in range(-n // 2, n // 2 + 1):
        x = flat.clone()
        x[:, j] += k * step
        with torch.no_grad():
            F.append((_fwd(c, dict(P, **{group: x.reshape(P[group].shape)})) * v).sum(dim=1).cpu().numpy())
    return np.median(np.diff(np.array(F), axis=0) / step, axis=0)


def test_vjp_batch64_vs_fd(vjp64):
    import torch
    c, draws, P, v = vjp64
    leaves = {k: x.clone().requires_grad_(True) for k, x in P.items()}
    P = dict(P, _v=v)
    (_fwd(c, leaves) * v).sum().backward()
    torch.cuda.synchronize()
    grads = {k: x.grad.detach().cpu().numpy() for k, x in leaves.items()}
    assert np.abs(grads["lnP"]).max() > 0
    B = len(draws)
    for group in ("src_dz", "lens_dz", "b1", "ia_A", "shear_m"):
        x0 = P[group]
        flat = x0.reshape(B, -1)
        fd = np.zeros(flat.shape)
        for j in range(flat.shape[1]):
            if group == "lens_dz":
                fd[:, j] = _median_slope(c, P, group, flat, j)
                continue
            h = 1e-4 * np.abs(flat[:, j].cpu().numpy())
            h = np.where(h > 0, h, 1e-5)
            ht = torch.tensor(h, dtype=torch.float64, device=x0.device)
            xp = flat.clone(); xp[:, j] += ht
            xm = flat.clone(); xm[:, j] -= ht
            with torch.no_grad():
                dp = _fwd(c, dict(P, **{group: xp.reshape(x0.shape)}))
                dm = _fwd(c, dict(P, **{group: xm.reshape(x0.shape)}))
                fd[:, j] = ((dp - dm) * v).sum(dim=1).cpu().numpy() / (2 * h)
        got = grads[group].reshape(fd.shape)
        floor = 1e-3 * np.abs(fd).max()
        err = np.abs(got - fd) / np.maximum(np.abs(fd), floor)
        print(group, "max rel err", err.max())
        assert err.max() < 1e-4, (group, err.max())
    u = torch.tensor(np.random.default_rng(9).normal(size=P["lnP"].shape), dtype=torch.float64,
                     device=P["lnP"].device)
    h = 1e-4
    with torch.no_grad():
        fd = ((_fwd(c, dict(P, lnP=P["lnP"] + h * u)) - _fwd(c, dict(P, lnP=P["lnP"] - h * u))) * v).sum(dim=1)
    fd = fd.cpu().numpy() / (2 * h)
    got = (grads["lnP"] * u.cpu().numpy()).sum(axis=1)
    err = np.abs(got - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    print("lnP direction max rel err", err.max())
    assert err.max() < 1e-4
