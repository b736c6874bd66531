"""Adjoint (VJP) of the batched FP64 path vs central finite differences.

Checked parameters: ln P(k,z) (random direction), source and lens photo-z
shifts, linear bias, A_IA and shear calibration m.  Finite differences use a
relative step of 1e-4 (absolute 1e-5 where the parameter is zero: the lens
integration limits move with |dz|, so at dz = 0 the central difference carries
an O(|h|) term -- 1.8e-4 relative at h = 1e-4, 8e-6 at h = 1e-5), of the GPU
forward itself for every scalar and of the FP64 oracle for a subset.  The
reference has no gradient code, so finite differences are the gradient pin.
Tolerance: |vjp - fd| <= 1e-4 max(|fd|, 1e-3 max|fd| of the parameter group).
"""
import numpy as np
import pytest

from cosmolike_amd import Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _draws():
    out = []
    for i, (cos, *_r) in enumerate(synth.random_draws(2, seed=5, ragged=False)):
        p = synth.full_3x2pt_problem(cos)
        rng = np.random.default_rng(100 + i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.b1 = np.linspace(1.2, 2.0, p.nl) * (1 + 0.05 * i)
        p.ia_A = 0.5 + 0.1 * i
        out.append(p)
    return out


@pytest.fixture(scope="module")
def setup():
    import torch
    draws = _draws()
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    params = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
                  lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
                  ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    rng = np.random.default_rng(42)
    v = t(rng.normal(size=(len(draws), c.datavector_size())))
    return c, draws, params, v


def _fwd(c, P):
    from cosmolike_amd.autograd import projection_3x2pt
    return projection_3x2pt(c, P["lnP"], P["src_dz"], P["lens_dz"], P["b1"], P["ia_A"], P["shear_m"])


def _vjp(c, P, v):
    import torch
    leaves = {k: x.clone().requires_grad_(True) for k, x in P.items()}
    out = _fwd(c, leaves)
    (out * v).sum().backward()
    torch.cuda.synchronize()
    return {k: x.grad.detach().cpu().numpy() for k, x in leaves.items()}


def _check(got, fd, name):
    got, fd = np.asarray(got), np.asarray(fd)
    floor = 1e-3 * np.abs(fd).max()
    err = np.abs(got - fd) / np.maximum(np.abs(fd), floor)
    print(name, "max rel err", err.max())
    assert err.max() < TOL, (name, got, fd)


def test_forward_matches_oracle(setup):
    import torch
    c, draws, P, v = setup
    out = _fwd(c, P).detach().cpu().numpy()
    for i, d in enumerate(draws):
        ref = O.Oracle().load(d).datavector()
        floor = 1e-3 * np.abs(ref).max()
        assert (np.abs(out[i] - ref) / np.maximum(np.abs(ref), floor)).max() < 1e-6


@pytest.mark.parametrize("group", ["src_dz", "lens_dz", "b1", "ia_A", "shear_m"])
def test_vjp_scalars_vs_gpu_fd(setup, group):
    import torch
    c, draws, P, v = setup
    g = _vjp(c, P, v)[group]
    x0 = P[group]
    flat = x0.reshape(len(draws), -1)
    fd = np.zeros(flat.shape)
    for j in range(flat.shape[1]):
        h = 1e-4 * np.maximum(np.abs(flat[:, j].cpu().numpy()), 0.0)
        h = np.where(h > 0, h, 1e-5)
        ht = torch.tensor(h, dtype=torch.float64, device=x0.device)
        Pp = dict(P); Pm = dict(P)
        xp = flat.clone(); xp[:, j] += ht
        xm = flat.clone(); xm[:, j] -= ht
        Pp[group] = xp.reshape(x0.shape); Pm[group] = xm.reshape(x0.shape)
        with torch.no_grad():
            d = ((_fwd(c, Pp) - _fwd(c, Pm)) * v).sum(dim=1).cpu().numpy()
        fd[:, j] = d / (2 * h)
    _check(g.reshape(fd.shape), fd, group)


def test_vjp_lnP_direction(setup):
    import torch
    c, draws, P, v = setup
    g = _vjp(c, P, v)["lnP"]
    rng = np.random.default_rng(9)
    u = torch.tensor(rng.normal(size=P["lnP"].shape), dtype=torch.float64, device=P["lnP"].device)
    h = 1e-4
    with torch.no_grad():
        Pp = dict(P, lnP=P["lnP"] + h * u)
        Pm = dict(P, lnP=P["lnP"] - h * u)
        fd = ((_fwd(c, Pp) - _fwd(c, Pm)) * v).sum(dim=1).cpu().numpy() / (2 * h)
    got = (g * u.cpu().numpy()).sum(axis=1)
    _check(got, fd, "lnP")


def test_vjp_vs_oracle_fd(setup):
    """Independent pin: FD of the FP64 CPU oracle for draw 0."""
    c, draws, P, v = setup
    g = _vjp(c, P, v)
    d0 = draws[0]
    vv = v[0].cpu().numpy()

    def f(**kw):
        o = O.Oracle().load(d0)
        o.set_nuisance(**kw)
        return float(vv @ o.datavector())

    base = dict(src_dz=d0.src_dz, lens_dz=d0.lens_dz, b1=d0.b1, shear_m=d0.shear_m)
    checks = []
    for name, j in (("src_dz", 3), ("src_dz", 8), ("lens_dz", 2), ("b1", 1), ("shear_m", 4)):
        x = np.array(base[name], dtype=float)
        h = 1e-4 * abs(x[j]) if x[j] != 0 else 1e-5
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        fd = (f(**dict(base, **{name: xp})) - f(**dict(base, **{name: xm}))) / (2 * h)
        checks.append((g[name][0][j], fd, f"{name}[{j}]"))
    for got, fd, name in checks:
        assert abs(got - fd) <= TOL * abs(fd), (name, got, fd)


def test_vjp_reuse_matches_recompute(setup):
    """A VJP right after an FP64 forward reuses its device state (node
    weights, C_ell rows); after a parameter update it recomputes them."""
    import torch
    c, draws, P, v = setup
    B, n = v.shape
    g = {k: torch.empty_like(x) for k, x in P.items()}
    ptrs = [g[k].data_ptr() for k in ("lnP", "src_dz", "lens_dz", "b1", "ia_A", "shear_m")]
    out = torch.empty(B * n, dtype=torch.float64, device="cuda:0")
    c.batch_set_params(*[P[k].data_ptr() for k in ("lnP", "src_dz", "lens_dz", "b1", "ia_A", "shear_m")])
    c.batch_forward(out.data_ptr())
    c.batch_vjp(v.data_ptr(), *ptrs)
    c.synchronize()
    torch.cuda.synchronize()
    reuse = {k: x.cpu().numpy().copy() for k, x in g.items()}
    c.batch_set_params(*[P[k].data_ptr() for k in ("lnP", "src_dz", "lens_dz", "b1", "ia_A", "shear_m")])
    c.batch_vjp(v.data_ptr(), *ptrs)
    c.synchronize()
    torch.cuda.synchronize()
    for k, x in g.items():
        r = x.cpu().numpy()
        assert np.abs(r - reuse[k]).max() <= 1e-11 * np.abs(r).max(), k


def test_vjp_checkpoint_matches_recompute(setup):
    """With clh_set_vjp_checkpoint the VJP reads the forward's gg integrand;
    it must equal the VJP that evaluates it again (up to the order of atomic
    sums), and a checkpoint taken before a parameter update must not be used."""
    import torch
    c, draws, P, v = setup
    B, n = v.shape
    keys = ("lnP", "src_dz", "lens_dz", "b1", "ia_A", "shear_m")
    g = {k: torch.empty_like(x) for k, x in P.items()}
    out = torch.empty(B * n, dtype=torch.float64, device="cuda:0")

    fwd = {}

    def run(PP, ckpt, update_after=None):
        c.set_vjp_checkpoint(ckpt)
        c.batch_set_params(*[PP[k].data_ptr() for k in keys])
        c.batch_forward(out.data_ptr())
        fwd[ckpt] = out.cpu().numpy().copy()
        if update_after is not None:
            c.batch_set_params(*[update_after[k].data_ptr() for k in keys])
        c.batch_vjp(v.data_ptr(), *[g[k].data_ptr() for k in keys])
        c.synchronize()
        torch.cuda.synchronize()
        return {k: x.cpu().numpy().copy() for k, x in g.items()}

    def same(a, b):
        for k in keys:
            assert np.abs(a[k] - b[k]).max() <= 1e-11 * np.abs(b[k]).max(), k

    P2 = dict(P, b1=P["b1"] * 1.1, lens_dz=P["lens_dz"] + 0.01)
    try:
        same(run(P, True), run(P, False))
        # the checkpointing forward (tables with tangents) gives the same data vector
        assert np.array_equal(fwd[True], fwd[False])
        same(run(P, True, update_after=P2), run(P2, False))
    finally:
        c.set_vjp_checkpoint(False)


def test_lens_limit_term_tangent_vs_views():
    """The lens-limit part of the lens-shift gradient: the GL nodes of a lens
    bin's gs and gg integrals move with |dz| (amin_lens / amax_lens,
    redshift_spline.c:50-79).  The default (clh_set_vjp_lens_limits 0) takes
    it from the nodes' motion inside the adjoint kernels; mode 1 from
    finite-difference views of the forward (|dz| +- 1e-8, per-row knot-jump
    rejection); mode 2 leaves it out.  The term (mode - mode 2) of modes 0 and
    1 agree to 2e-3 of its size, on draws with lens shifts of both signs,
    stretched lens bins and magnification bias (amax then fixed); the other
    gradients do not depend on the mode."""
    import torch
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(3, seed=11, ragged=False)):
        p = synth.full_3x2pt_problem(cos)
        rng = np.random.default_rng(300 + i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.02, p.nl)
        p.lens_stretch = 1 + rng.normal(0, 0.05, p.nl)
        p.b1 = np.linspace(1.2, 2.0, p.nl) * (1 + 0.05 * i)
        if i == 1:
            p.bmag = rng.normal(0, 1.0, p.nl)
        draws.append(p)
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(43).normal(size=(len(draws), c.datavector_size())))
    g = {}
    try:
        for mode in (0, 1, 2):
            c.set_vjp_lens_limits(mode)
            g[mode] = _vjp(c, P, v)
    finally:
        c.set_vjp_lens_limits(0)
    term0 = g[0]["lens_dz"] - g[2]["lens_dz"]
    term1 = g[1]["lens_dz"] - g[2]["lens_dz"]
    size = np.abs(term1).max(axis=1)
    scale = np.abs(g[1]["lens_dz"]).max(axis=1)
    err = np.abs(term0 - term1).max(axis=1) / size
    print("lens-limit term / gradient scale", size / scale, "tangent vs views", err)
    assert (size > 1e-4 * scale).all()
    assert err.max() < 2e-3, (term0, term1)
    for k in ("lnP", "src_dz", "b1", "ia_A", "shear_m"):
        a, b = g[0][k], g[2][k]
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max(), k


def _median_slope(c, P, v, group, flat, j, step=2e-8, n=8):
    """Median over k of (F(x + (k+1) step) - F(x + k step)) / step with
    F = v . forward, per draw: forward slopes that leave out a step holding a
    table-knot jump of the integrands."""
    import torch
    F = []
    for k in range(-n // 2, n // 2 + 1):
        x = flat.clone()
        x[:, j] += k * step
        with torch.no_grad():
            F.append((_fwd(c, dict(P, **{group: x.reshape(P[group].shape)})) * v).sum(dim=1).cpu().numpy())
    return np.median(np.diff(np.array(F), axis=0) / step, axis=0)


def test_vjp_oneloop_vs_gpu_fd():
    """The VJP with one-loop galaxy bias (b2, bs2, b3, bK; FAST-PT bias
    tables) against differences of the GPU forward: every scalar group and a
    random ln P direction at 1e-4 (as the NLA groups), lens shifts of both
    signs with magnification bias on one draw (the one-loop batches take the
    lens-limit part from the finite-difference views).  d/d b1 is taken at
    the uploaded bs2, as the device-side update of b1 leaves bs2."""
    import torch
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(2, seed=23, ragged=False)):
        p = synth.oneloop_problem(cos)
        rng = np.random.default_rng(500 + i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.005, p.nl)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.ia_A = 0.5 + 0.1 * i
        if i == 1:
            p.bmag = rng.normal(0, 1.0, p.nl)
        draws.append(p)
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(44).normal(size=(len(draws), c.datavector_size())))
    # the forward with the one-loop terms against the oracle
    out = _fwd(c, P).detach().cpu().numpy()
    for i, d in enumerate(draws):
        ref = O.Oracle().load(d).datavector()
        floor = 1e-3 * np.abs(ref).max()
        assert (np.abs(out[i] - ref) / np.maximum(np.abs(ref), floor)).max() < 1e-6
    _groups_vs_fd(c, P, v, len(draws), "oneloop")


def _groups_vs_fd(c, P, v, B, tag):
    """Every scalar group and a random ln P direction of the VJP against
    differences of the GPU forward (relative step 1e-4, absolute 1e-5 at zero;
    the lens shifts by median slopes)."""
    import torch
    g = _vjp(c, P, v)
    for group in ("src_dz", "lens_dz", "b1", "ia_A", "shear_m"):
        x0 = P[group]
        flat = x0.reshape(B, -1)
        fd = np.zeros(flat.shape)
        for j in range(flat.shape[1]):
            if group == "lens_dz":
                fd[:, j] = _median_slope(c, P, v, group, flat, j)
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
        _check(g[group].reshape(fd.shape), fd, tag + " " + group)
    _lnP_dir_vs_fd(c, P, v, g, tag)


def _lnP_dir_vs_fd(c, P, v, g, tag, h=1e-4):
    """The ln P gradient g along a random direction against the central
    difference of v . forward (step h)."""
    import torch
    rng = np.random.default_rng(10)
    u = torch.tensor(rng.normal(size=P["lnP"].shape), dtype=torch.float64, device=P["lnP"].device)
    with torch.no_grad():
        fd = ((_fwd(c, dict(P, lnP=P["lnP"] + h * u)) - _fwd(c, dict(P, lnP=P["lnP"] - h * u))) * v).sum(dim=1)
    _check((g["lnP"] * u.cpu().numpy()).sum(axis=1), fd.cpu().numpy() / (2 * h), tag + " lnP")


def test_vjp_tatt_vs_gpu_fd():
    """The VJP under TATT intrinsic alignments (A1 = ia_A with redshift
    evolution, A2, eta2 and b_TA fixed; FAST-PT IA tables as inputs): the EE
    and BB C_ss rows and the TATT C_gs rows back through the node vectors
    A = W_K - C1 W_S, B = W_S C1 b_TA, D = W_S C2 to the source shifts and
    A_IA, and through P to ln P.  Every scalar group and a ln P direction
    against differences of the GPU forward at 1e-4; lens shifts of both signs
    with magnification bias on one draw (the lens-limit part from the
    finite-difference views)."""
    import torch
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(2, seed=29, ragged=False)):
        p = synth.tatt_problem(cos, A2=-0.8 + 0.3 * i, eta2=0.5, bta=0.7 - 0.2 * i)
        rng = np.random.default_rng(700 + i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.005, p.nl)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.ia_A = 0.5 + 0.2 * i
        if i == 1:
            p.bmag = rng.normal(0, 1.0, p.nl)
        draws.append(p)
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(45).normal(size=(len(draws), c.datavector_size())))
    out = _fwd(c, P).detach().cpu().numpy()
    for i, d in enumerate(draws):
        ref = O.Oracle().load(d).datavector()
        floor = 1e-3 * np.abs(ref).max()
        assert (np.abs(out[i] - ref) / np.maximum(np.abs(ref), floor)).max() < 1e-6
    _groups_vs_fd(c, P, v, len(draws), "tatt")


def _sixpt_draws(seed, bp=None):
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(2, seed=seed, ragged=False)):
        p = synth.sixpt_problem(cos)
        rng = np.random.default_rng(seed + 10 * i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.005, p.nl)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.ia_A = 0.5 + 0.2 * i
        if i == 1:
            p.bmag = rng.normal(0, 1.0, p.nl)
        if bp is not None:
            p.kk_bp = bp
        draws.append(p)
    return draws


def _sixpt_setup(draws, mask_seed=None):
    import torch
    c = Cosmolike(device=0).load(draws[0])
    if mask_seed is not None:
        n = c.datavector_size()
        mask = (np.random.default_rng(mask_seed).random(n) > 0.2).astype(np.int32)
        c.like_set_mask(mask)
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(46).normal(size=(len(draws), c.datavector_size())))
    return c, P, v


def test_vjp_6x2pt_vs_gpu_fd():
    """The VJP of the 6x2pt data vector: the w_gk, w_ks and C_kk blocks back
    through the CMB-filtered kernels and C_kk(ell) to the C_gk / C_ks / C_kk
    rows, and through their Limber integrands to ln P, the lens shifts (W_gal,
    W_mag; the gk limits' motion from the finite-difference views), b1, the
    source shifts and A_IA (C_ks) and m (w_ks).  The forward against the
    oracle first; then every scalar group and a ln P direction against
    differences of the GPU forward at 1e-4, magnification bias on one draw."""
    draws = _sixpt_draws(31)
    c, P, v = _sixpt_setup(draws)
    ells = synth.ells_fourier(draws[0].kk_nells, draws[0].kk_lmin, draws[0].kk_lmax)
    out = _fwd(c, P).detach().cpu().numpy()
    for i, d in enumerate(draws):
        ref = O.Oracle().load(d).datavector_6x2pt(ells)
        floor = 1e-3 * np.abs(ref).max()
        assert (np.abs(out[i] - ref) / np.maximum(np.abs(ref), floor)).max() < 1e-6
    _groups_vs_fd(c, P, v, len(draws), "6x2pt")
    # cotangents on the CMB blocks alone (w_gk, w_ks, C_kk after the 3x2pt
    # blocks), and on C_kk alone (ln P only), so that the small C_kk entries
    # are not hidden behind the others.  The C_kk integral runs to z ~ 1000,
    # far past the P(k,z) table (z <= 3), where the bilinear z weights
    # extrapolate with |dz| in the hundreds: a random ln P step of 1e-4 moves
    # ln P there by 0.01-0.1 and the central difference carries that
    # curvature, so its step is 1e-6.
    p = draws[0]
    n3 = (p.ns * (p.ns + 1) + p.ns * p.nl + p.nl) * p.ntheta
    v6 = v.clone()
    v6[:, :n3] = 0.0
    _groups_vs_fd(c, P, v6, len(draws), "6x2pt CMB blocks")
    vkk = v.clone()
    vkk[:, :n3 + (p.nl + p.ns) * p.ntheta] = 0.0
    _lnP_dir_vs_fd(c, P, vkk, _vjp(c, P, vkk), "6x2pt kk", h=1e-6)


def test_vjp_6x2pt_bandpowers_mask_vs_gpu_fd():
    """As above with kk band powers (init_cmb_auto_bandpower: the cotangent of
    every band spread over C_kk(L), L = lmin..lmax, through the direct rows
    below LMIN_tab and the table rows above) and a mask over every block
    (masked entries carry no cotangent)."""
    draws = _sixpt_draws(37, bp=synth.kk_bandpower_binning())
    c, P, v = _sixpt_setup(draws, mask_seed=8)
    _groups_vs_fd(c, P, v, len(draws), "6x2pt bp")
    nbp = draws[0].kk_bp[0]
    vkk = v.clone()
    vkk[:, :-nbp] = 0.0
    _lnP_dir_vs_fd(c, P, vkk, _vjp(c, P, vkk), "6x2pt bp kk", h=1e-6)


def test_vjp_nonlimber_vs_gpu_fd():
    """The VJP with the non-Limber w(theta) below ell = 150: the w cotangent
    onto the FFTLog C_ell (its sums sum_q F^2 k^3 P_lin are linear in each
    transformed input, differentiated by tangent transforms of the inputs'
    b1 and lens-shift derivatives), onto the integer-ell Limber rows with P_nl
    (ln P, b1, lens shift) and P_lin (b1, lens shift), and off the Limber
    C_ell the folded kernel summed.  The forward against the oracle first;
    then every group and a ln P direction against differences of the GPU
    forward at 1e-4 with cotangents on the whole vector, on the w block alone
    and on its entries above 100 arcmin (where the correction is largest),
    magnification bias on one draw."""
    import torch
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(2, seed=41, ragged=False)):
        p = synth.nonlimber_problem(cos)
        rng = np.random.default_rng(900 + i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.005, p.nl)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.b1 = np.linspace(1.2, 2.0, p.nl) * (1 + 0.1 * i)
        p.ia_A = 0.5 + 0.2 * i
        if i == 1:
            p.bmag = rng.normal(0, 1.0, p.nl)
        draws.append(p)
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(47).normal(size=(len(draws), c.datavector_size())))
    out = _fwd(c, P).detach().cpu().numpy()
    for i, d in enumerate(draws):
        ref = O.Oracle().load(d).datavector()
        floor = 1e-3 * np.abs(ref).max()
        assert (np.abs(out[i] - ref) / np.maximum(np.abs(ref), floor)).max() < 1e-6
    _groups_vs_fd(c, P, v, len(draws), "nonlimber")
    p = draws[0]
    nw0 = (p.ns * (p.ns + 1) + p.ns * p.nl) * p.ntheta  # the w block's start
    vw = torch.zeros_like(v)
    vw[:, nw0:nw0 + p.nl * p.ntheta] = v[:, nw0:nw0 + p.nl * p.ntheta]
    _w_groups_vs_fd(c, P, vw, len(draws), "nonlimber w")
    # the largest third of the log-spaced angles (2.5'-900': above ~100')
    vb = torch.zeros_like(v)
    for zl in range(p.nl):
        o = nw0 + zl * p.ntheta
        k0 = o + p.ntheta - p.ntheta // 3
        vb[:, k0:o + p.ntheta] = v[:, k0:o + p.ntheta]
    _w_groups_vs_fd(c, P, vb, len(draws), "nonlimber w large theta")


def _w_groups_vs_fd(c, P, v, B, tag):
    """The groups a w-block cotangent reaches (ln P, lens shifts, b1)."""
    import torch
    g = _vjp(c, P, v)
    for group in ("lens_dz", "b1"):
        x0 = P[group]
        flat = x0.reshape(B, -1)
        fd = np.zeros(flat.shape)
        for j in range(flat.shape[1]):
            if group == "lens_dz":
                fd[:, j] = _median_slope(c, P, v, group, flat, j)
                continue
            h = 1e-4 * np.abs(flat[:, j].cpu().numpy())
            ht = torch.tensor(h, dtype=torch.float64, device=x0.device)
            xp = flat.clone(); xp[:, j] += ht
            xm = flat.clone(); xm[:, j] -= ht
            with torch.no_grad():
                dp = _fwd(c, dict(P, **{group: xp.reshape(x0.shape)}))
                dm = _fwd(c, dict(P, **{group: xm.reshape(x0.shape)}))
            fd[:, j] = ((dp - dm) * v).sum(dim=1).cpu().numpy() / (2 * h)
        _check(g[group].reshape(fd.shape), fd, tag + " " + group)
    _lnP_dir_vs_fd(c, P, v, g, tag)


@pytest.mark.parametrize("hdi", [1, -2])
def test_vjp_hdi_vs_gpu_fd(hdi):
    """The VJP at high_def_integration != 0 (NLA): the adjoint of the n-point
    GL rows (k_limber_hdi_adj: ln P runs per node, the radial weights'
    shift, b1 and A_IA tangents at each node) and the gs / gg limits' motion
    with |dz| from the finite-difference views, which rerun k_limber_hdi.
    The forward against the oracle at the same level first; then every group
    and a ln P direction against differences of the GPU forward at 1e-4,
    lens stretch != 1 and magnification bias on one draw."""
    import torch
    draws = []
    for i, (cos, *_r) in enumerate(synth.random_draws(2, seed=53, ragged=False)):
        p = synth.full_3x2pt_problem(cos)
        rng = np.random.default_rng(1100 + i)
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.lens_dz = rng.normal(0, 0.005, p.nl)
        p.lens_stretch = 1.0 + rng.normal(0, 0.02, p.nl)
        p.shear_m = rng.normal(0, 0.01, p.ns)
        p.b1 = np.linspace(1.2, 2.0, p.nl) * (1 + 0.1 * i)
        p.ia_A = 0.5 + 0.2 * i
        if i == 1:
            p.bmag = rng.normal(0, 1.0, p.nl)
        p.hdi = hdi
        draws.append(p)
    c = Cosmolike(device=0).load(draws[0])
    c.batch_upload(draws)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.array(a), dtype=torch.float64, device=dev)
    P = dict(lnP=t([d.lnP for d in draws]), src_dz=t([d.src_dz for d in draws]),
             lens_dz=t([d.lens_dz for d in draws]), b1=t([d.b1 for d in draws]),
             ia_A=t([d.ia_A for d in draws]), shear_m=t([d.shear_m for d in draws]))
    v = t(np.random.default_rng(48).normal(size=(len(draws), c.datavector_size())))
    out = _fwd(c, P).detach().cpu().numpy()
    for i, d in enumerate(draws):
        ref = O.Oracle().load(d).datavector()
        floor = 1e-3 * np.abs(ref).max()
        assert (np.abs(out[i] - ref) / np.maximum(np.abs(ref), floor)).max() < 1e-6
    _groups_vs_fd(c, P, v, len(draws), "hdi %d" % hdi)
