"""Batches whose draws differ in survey layout (3-10 source bins, 2-6 lens
bins, 20-30 theta bins) in one set of launches, vs the FP64 CPU oracle run
draw by draw with each draw's own layout.

Bar: FP64 path 1e-6 relative (metric as in the FP64 parity tests); the graph
replay reproduces the eager launches bit for bit.
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


@pytest.fixture(scope="module")
def ragged():
    spec = synth.random_draws(10, seed=21, ragged=True)
    probs = [synth.ragged_3x2pt_problem(c, ns, nl, nt) for c, ns, nl, nt in spec]
    rng = np.random.default_rng(3)
    for p in probs:
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.shear_m = rng.normal(0, 0.01, p.ns)
    c = Cosmolike()
    cache, lays = {}, []
    for p in probs:
        key = (p.ns, p.nl, p.ntheta)
        if key not in cache:
            c.load(p)
            cache[key] = c.capture_layout()
        lays.append(cache[key])
    refs = [O.Oracle().load(p).datavector() for p in probs]
    return c, probs, lays, refs


def _run(c, n):
    import torch
    out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_ragged_fp64_matches_oracle(ragged):
    c, probs, lays, refs = ragged
    c.set_precision("fp64")
    c.batch_upload_layouts(probs, lays)
    offs = c.batch_offsets()
    assert offs[-1] == sum(l.datavector_size() for l in lays)
    got = _run(c, offs[-1])
    for i, p in enumerate(probs):
        e = rel_err(got[offs[i]:offs[i + 1]], refs[i], p.ns, p.nl, p.ntheta)
        print(f"draw {i} ns={p.ns} nl={p.nl} nt={p.ntheta} err {e:.2e}")
        assert e < 1e-6


def test_ragged_graph_replay(ragged):
    c, probs, lays, refs = ragged
    c.set_precision("fp64")
    c.batch_upload_layouts(probs, lays)
    n = c.batch_offsets()[-1]
    eager = _run(c, n)
    c.batch_use_graph(True)
    try:
        for _ in range(3):
            assert np.array_equal(_run(c, n), eager)
    finally:
        c.batch_use_graph(False)


@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
def test_ragged_mixed_precision(ragged, mode):
    c, probs, lays, refs = ragged
    c.set_precision(mode)
    try:
        c.batch_upload_layouts(probs, lays)
        offs = c.batch_offsets()
        got = _run(c, offs[-1])
    finally:
        c.set_precision("fp64")
    worst = max(rel_err(got[offs[i]:offs[i + 1]], refs[i], p.ns, p.nl, p.ntheta) for i, p in enumerate(probs))
    print(mode, "worst", worst)
    assert worst <= 1e-4


def test_ragged_batch_refuses_vjp(ragged):
    import torch
    c, probs, lays, refs = ragged
    c.batch_upload_layouts(probs, lays)
    n = c.batch_offsets()[-1]
    dvbar = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    with pytest.raises(CosmolikeError) as e:
        c.batch_vjp(dvbar.data_ptr())
    assert e.value.code == -1
