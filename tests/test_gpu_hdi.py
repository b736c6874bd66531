"""high_def_integration != 0 on the GPU against the oracle at the same level.

At |hdi| >= 1 the Limber rows are n-point GL sums of the integrands
(k_limber_hdi: 96 .. 512 nodes for ss / gs, 128 .. 1024 for gg), the lens
efficiencies use 25 (1 + |hdi|) fine-grid points per interval and the lens
zmean a 512 / 1024-point rule.  Bar: 1e-6 relative (metric of the FP64 parity
tests), 1e-4 in the FP32 mode; graph replays bit-identical; the 6x2pt
probes and the non-Limber w refuse hdi != 0, and so does the VJP under TATT
(the VJP of NLA draws is in test_gpu_vjp.py).
"""
import numpy as np
import pytest

from cosmolike_amd import Cosmolike, CosmolikeError, synth

import oracle as O

pytestmark = pytest.mark.gpu


def rel_err(gpu, ref, ns, nl, nt):
    n_ss, n_gs = ns * (ns + 1) // 2, ns * nl
    worst, o = 0.0, 0
    for rows in (n_ss, n_ss, n_gs, nl):
        g = gpu[o:o + rows * nt].reshape(rows, nt)
        r = ref[o:o + rows * nt].reshape(rows, nt)
        floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
        worst = max(worst, float((np.abs(g - r) / np.maximum(np.abs(r), floor)).max()))
        o += rows * nt
    return worst


def tab_err(g, r):
    floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(r), floor)
    den[den == 0] = 1.0
    return float((np.abs(g - r) / den).max())


def _c2(hdi, seed=5):
    cos = synth.random_draws(1, seed=seed, ragged=False)[0][0]
    p = synth.full_3x2pt_problem(cos)
    rng = np.random.default_rng(seed)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.lens_stretch = 1.0 + rng.normal(0, 0.02, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    p.bmag = np.array([0.0, 0.9, 0.0, 0.0, -0.4])
    p.hdi = hdi
    return p


@pytest.mark.parametrize("hdi", [1, 2, -4])
def test_hdi_datavector_and_tables(hdi):
    p = _c2(hdi)
    print("hdi", hdi, flush=True)
    c = Cosmolike().load(p)
    o = O.Oracle().load(p)
    e = rel_err(c.compute_3x2pt_real(), o.datavector(), p.ns, p.nl, p.ntheta)
    ss, gs, gg = c.cl_tables(p.n_ell)
    rs, rg, rgg = o.cl_tables()
    et = max(tab_err(ss, rs), tab_err(gs, rg), tab_err(gg, rgg))
    print("hdi", hdi, "dv err", e, "table err", et)
    assert e < 1e-6 and et < 1e-6
    # the level changes the result (the tables are not the hdi = 0 ones)
    p0 = _c2(0)
    d0 = O.Oracle().load(p0).datavector()
    assert np.abs(o.datavector() - d0).max() > 1e-9 * np.abs(d0).max()


@pytest.mark.parametrize("model", ["tatt", "oneloop"])
def test_hdi_models(model):
    cos = synth.random_draws(1, seed=31, ragged=False)[0][0]
    p = synth.tatt_problem(cos) if model == "tatt" else synth.oneloop_problem(cos)
    p.hdi = 1
    print("model", model, flush=True)
    c = Cosmolike().load(p)
    o = O.Oracle().load(p)
    e = rel_err(c.compute_3x2pt_real(), o.datavector(), p.ns, p.nl, p.ntheta)
    print(model, "err", e)
    assert e < 1e-6


def test_hdi_fourier_and_fp32():
    p = _c2(1, seed=9)
    c = Cosmolike().load(p)
    c.init_binning_fourier(12, 30, 3000, 2000)
    o = O.Oracle().load(p)
    ells = np.exp(np.linspace(np.log(30.0), np.log(3000.0), 13))
    ells = np.sqrt(ells[1:] * ells[:-1])
    ref = o.datavector_fourier(ells, 2000, (0, 1, 2))
    got = c.compute_3x2pt_fourier((0, 1, 2))
    floor = 1e-3 * np.abs(ref).max()
    e = float((np.abs(got - ref) / np.maximum(np.abs(ref), floor)).max())
    print("fourier err", e)
    assert e < 1e-6
    ref = o.datavector()
    try:
        c.set_precision("fp32")
        e32 = rel_err(c.compute_3x2pt_real(), ref, p.ns, p.nl, p.ntheta)
    finally:
        c.set_precision("fp64")
    print("fp32 err", e32)
    assert e32 < 1e-4


def test_hdi_mixed_levels_batch_graph():
    """Draws at hdi 0 and 3 in one batch of captured layouts (the hdi = 0 draws
    in the node kernels, the others in k_limber_hdi), every draw against the
    oracle, graph replays bit-identical to the eager launches."""
    import torch
    probs = [_c2(h, seed=40 + i) for i, h in enumerate((0, 3, 0, 3))]
    c = Cosmolike()
    lays = []
    for p in probs:
        c.load(p)
        lays.append(c.capture_layout())
    refs = [O.Oracle().load(p).datavector() for p in probs]
    c.batch_upload_layouts(probs, lays)
    offs = c.batch_offsets()
    n = offs[-1]

    def run():
        out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0")
        c.batch_forward(out.data_ptr())
        c.synchronize()
        torch.cuda.synchronize()
        return out.cpu().numpy()

    eager = run()
    for i, p in enumerate(probs):
        e = rel_err(eager[offs[i]:offs[i + 1]], refs[i], p.ns, p.nl, p.ntheta)
        print("draw", i, "hdi", p.hdi, "err", e)
        assert e < 1e-6
    c.batch_use_graph(True)
    try:
        for _ in range(3):
            assert np.array_equal(run(), eager)
    finally:
        c.batch_use_graph(False)


def test_hdi_refusals():
    import torch
    p = synth.tatt_problem()
    p.hdi = 1
    c = Cosmolike().load(p)
    c.batch_upload([p])
    n = c.datavector_size()
    dv = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    c.batch_forward(dv.data_ptr())
    g = torch.zeros(p.nl, dtype=torch.float64, device="cuda:0")
    with pytest.raises(CosmolikeError, match="high_def_integration"):
        c.batch_vjp(dv.data_ptr(), g_lens_dz=g.data_ptr())
    nl = synth.nonlimber_problem()
    nl.hdi = 2
    with pytest.raises(CosmolikeError, match="high_def_integration"):
        Cosmolike().load(nl).compute_3x2pt_real()
    six = synth.sixpt_problem()
    six.hdi = 2
    with pytest.raises(CosmolikeError, match="high_def_integration"):
        Cosmolike().load(six).compute_6x2pt_real()
