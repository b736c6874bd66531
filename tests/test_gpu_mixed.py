"""Mixed-precision paths (FP32 Limber + FP32 or 3-term split-BF16 MFMA
Legendre) vs the FP64 CPU oracle, per probe.

Error metric per entry: |gpu - ref| / max(|ref|, 1e-3 max|ref| of its
(probe, pair) row), the same metric as the FP64 parity tests.  Bar: max over
every entry of every probe <= 1e-4; the max and RMS per probe are printed.
"""
import numpy as np
import pytest

from cosmolike_amd import Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-4
MODES = ["fp32", "bf16x3"]


def probe_errors(dv, ref, ns, nl, nt):
    n_ss, n_gs = ns * (ns + 1) // 2, ns * nl
    sizes = {"xi+": n_ss, "xi-": n_ss, "gammat": n_gs, "w": nl}
    out, o = {}, 0
    for name, rows in sizes.items():
        if rows == 0:
            continue
        g = dv[o:o + rows * nt].reshape(rows, nt)
        r = ref[o:o + rows * nt].reshape(rows, nt)
        floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
        e = np.abs(g - r) / np.maximum(np.abs(r), np.where(floor > 0, floor, 1.0))
        out[name] = (float(e.max()), float(np.sqrt((e ** 2).mean())))
        o += rows * nt
    return out


def _report(tag, errs):
    for k, (mx, rms) in errs.items():
        print(f"{tag} {k:7s} max {mx:.3e} rms {rms:.3e}")


@pytest.fixture(scope="module")
def full3x2():
    p = synth.full_3x2pt_problem()
    rng = np.random.default_rng(11)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    return p, O.Oracle().load(p).datavector()


@pytest.mark.parametrize("mode", MODES)
def test_full_3x2pt_datavector_mixed(full3x2, mode):
    p, ref = full3x2
    c = Cosmolike().load(p)
    c.set_precision(mode)
    dv = c.compute_3x2pt_real()
    errs = probe_errors(dv, ref, p.ns, p.nl, p.ntheta)
    _report(mode, errs)
    assert max(v[0] for v in errs.values()) <= TOL, errs


@pytest.mark.parametrize("mode", MODES)
def test_shear_only_mixed(mode):
    p = synth.shear_only_problem()
    o = O.Oracle().load(p)
    xp, xm = o.xi_pm()
    c = Cosmolike().load(p)
    c.set_precision(mode)
    gp, gm = c.xi_pm()
    errs = probe_errors(np.concatenate([gp.ravel(), gm.ravel()]),
                        np.concatenate([xp.ravel(), xm.ravel()]), p.ns, 0, p.ntheta)
    _report("shear-only " + mode, errs)
    assert max(v[0] for v in errs.values()) <= TOL, errs


@pytest.mark.parametrize("mode", MODES)
def test_batch_mixed_random_cosmologies(mode):
    import torch
    draws = [synth.full_3x2pt_problem(cos) for cos, *_ in synth.random_draws(4, seed=3, ragged=False)]
    c = Cosmolike().load(draws[0])
    c.set_precision(mode)
    c.batch_upload(draws)
    n = c.datavector_size()
    out = torch.zeros((len(draws), n), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    worst = 0.0
    for i, d in enumerate(draws):
        ref = O.Oracle().load(d).datavector()
        errs = probe_errors(got[i], ref, d.ns, d.nl, d.ntheta)
        _report(f"draw{i} {mode}", errs)
        worst = max(worst, max(v[0] for v in errs.values()))
    assert worst <= TOL


def test_precision_switch_back_to_fp64(full3x2):
    p, ref = full3x2
    c = Cosmolike().load(p)
    c.set_precision("bf16x3")
    c.compute_3x2pt_real()
    c.set_precision("fp64")
    dv = c.compute_3x2pt_real()
    errs = probe_errors(dv, ref, p.ns, p.nl, p.ntheta)
    assert max(v[0] for v in errs.values()) < 1e-6, errs


@pytest.mark.parametrize("ckpt", [False, True])
@pytest.mark.parametrize("mode", MODES)
def test_vjp_after_mixed_forward(mode, ckpt):
    """The VJP in a mixed-precision mode: lens efficiencies and node weights
    are FP64 in every mode and the ss / gs adjoints evaluate in FP64, while
    the gg adjoint evaluates its integrand in FP32 (as the FP32 forward does)
    with FP64 cotangent sums; d/dm reads the ss C_ell rows of the forward
    through T.  With the integrand checkpoint the VJP reads the forward's FP32
    integrand instead.  Every gradient agrees with the FP64 mode's to 1e-5 of
    its largest entry."""
    import torch
    draws = [synth.full_3x2pt_problem(cos) for cos, *_ in synth.random_draws(3, seed=12, ragged=False)]
    c = Cosmolike().load(draws[0])
    c.batch_upload(draws)
    B, n, p = len(draws), c.datavector_size(), draws[0]
    dev = torch.device("cuda:0")
    out = torch.empty(B * n, dtype=torch.float64, device=dev)
    dvbar = torch.linspace(-1.0, 1.0, B * n, dtype=torch.float64, device=dev)
    sizes = (("lnP", p.lnP.size), ("src", p.ns), ("lens", p.nl), ("b1", p.nl), ("A", 1), ("m", p.ns))

    def grads(prec):
        c.set_precision(prec)
        c.batch_forward(out.data_ptr())
        g = {k: torch.empty(B * m, dtype=torch.float64, device=dev) for k, m in sizes}
        c.batch_vjp(dvbar.data_ptr(), *[g[k].data_ptr() for k, _ in sizes])
        c.synchronize()
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in g.items()}

    try:
        ref = grads("fp64")
        c.set_vjp_checkpoint(ckpt)
        got = grads(mode)
    finally:
        c.set_precision("fp64")
        c.set_vjp_checkpoint(False)
    tol = 1e-5
    for k in ("lnP", "src", "lens", "b1", "A"):
        e = np.abs(got[k] - ref[k]).max() / np.abs(ref[k]).max()
        print(mode, ckpt, k, e)
        assert e <= tol, k
    e = np.abs(got["m"] - ref["m"]).max() / np.abs(ref["m"]).max()
    print(mode, "d/dm rel", e)
    assert e <= 1e-5
