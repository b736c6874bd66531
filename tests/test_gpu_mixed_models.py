"""One batch whose draws use different models through per-draw captured
layouts: NLA 3x2pt, TATT intrinsic alignments, the one-loop galaxy bias,
non-Limber w, 6x2pt, and 6x2pt with the non-Limber w and kk band powers,
launched together, each draw against the FP64 oracle
with its own layout.

Bar: 1e-6 relative per entry with a floor of 1e-3 of the block row's largest
magnitude (the FP64 parity metric); the split-precision modes 1e-4 (FP32 and
BF16X3), their bars elsewhere.
"""
import numpy as np
import pytest

from cosmolike_amd import Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-6


def rel_err(gpu, ref, rows, width):
    g = gpu.reshape(rows, width)
    r = ref.reshape(rows, width)
    floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(r), floor)
    den[den == 0] = 1.0
    return float((np.abs(g - r) / den).max())


def block_errs(dv, ref, p, nf=0):
    nt, nss, ngs = p.ntheta, p.ns * (p.ns + 1) // 2, p.ns * p.nl
    spec = [("xi+", nss, nt), ("xi-", nss, nt), ("gammat", ngs, nt), ("w", p.nl, nt)]
    if nf:
        spec += [("gk", p.nl, nt), ("ks", p.ns, nt), ("kk", 1, nf)]
    out, o = {}, 0
    for name, rows, width in spec:
        out[name] = rel_err(dv[o:o + rows * width], ref[o:o + rows * width], rows, width)
        o += rows * width
    assert o == len(dv) == len(ref)
    return out


@pytest.mark.parametrize("mode,tol", [("fp64", RTOL), ("fp32", 1e-4), ("bf16x3", 1e-4)])
def test_models_in_one_batch(mode, tol):
    import torch
    cos = [c for c, *_ in synth.random_draws(6, seed=61, ragged=False)]
    probs = [synth.full_3x2pt_problem(cos[0]), synth.tatt_problem(cos[1]), synth.oneloop_problem(cos[2]),
             synth.nonlimber_problem(cos[3]), synth.sixpt_problem(cos[4]), synth.sixpt_problem(cos[5])]
    six_nl = probs[5]
    six_nl.plin_log10k, six_nl.plin_z, six_nl.lnPL = synth.linear_power_table(cos[5])
    six_nl.limber_gg = 0
    six_nl.kk_bp = synth.kk_bandpower_binning()
    rng = np.random.default_rng(61)
    for p in probs:
        p.src_dz = rng.normal(0, 0.01, p.ns)
        p.shear_m = rng.normal(0, 0.01, p.ns)
    refs, nfs = [], []
    for p in probs:
        o = O.Oracle().load(p)
        if getattr(p, "sixpt", 0):
            ells = synth.ells_fourier(p.kk_nells, p.kk_lmin, p.kk_lmax)
            refs.append(o.datavector_6x2pt(ells))
            nfs.append(p.kk_bp[0] if getattr(p, "kk_bp", None) is not None else len(ells))
        else:
            refs.append(o.datavector())
            nfs.append(0)
    c = Cosmolike()
    lays = []
    for p in probs:
        c.load(p)
        lays.append(c.capture_layout())
    order = [0, 1, 2, 3, 4, 5, 0]
    c.set_precision(mode)
    c.batch_upload_layouts([probs[k] for k in order], [lays[k] for k in order])
    offs = c.batch_offsets()
    assert offs[-1] == sum(len(refs[k]) for k in order)
    out = torch.full((offs[-1],), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for i, k in enumerate(order):
        e = block_errs(got[offs[i]:offs[i + 1]], refs[k], probs[k], nfs[k])
        print(i, k, e)
        assert max(e.values()) < tol, (i, k)
    # the same batch as one graph, replayed: bit for bit the eager result
    c.batch_use_graph(True)
    try:
        for _ in range(2):
            out.fill_(float("nan"))
            c.batch_forward(out.data_ptr())
            c.synchronize()
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy(), got)
    finally:
        c.batch_use_graph(False)
        c.set_precision("fp64")
