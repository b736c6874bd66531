"""6x2pt (init_probes("6x2pt")) on the GPU vs the FP64 oracle: the CMB lensing
probes w_gk(theta), w_ks(theta) (with the CMB beam and pixel filter) and
C_kk at the Fourier ell bins, the whole six-block real-space data vector in
two block orders and both Legendre modes, batches with graph replay, FP32
mode, kk band powers, masks, batches of different 6x2pt layouts, and the
switch back to 3x2pt (the VJP is in test_gpu_vjp.py).

Bar: 1e-6 relative per entry with a floor of 1e-3 of the block row's largest
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


def blocks(dv, p, nf):
    nt, nss, ngs, nl, ns = p.ntheta, p.ns * (p.ns + 1) // 2, p.ns * p.nl, p.nl, p.ns
    out, o = {}, 0
    for name, rows, width in (("xi+", nss, nt), ("xi-", nss, nt), ("gammat", ngs, nt), ("w", nl, nt),
                              ("gk", nl, nt), ("ks", ns, nt), ("kk", 1, nf)):
        out[name] = dv[o:o + rows * width].reshape(rows, width)
        o += rows * width
    assert o == len(dv)
    return out


def dv_err(dv, ref, p, nf):
    g, r = blocks(dv, p, nf), blocks(ref, p, nf)
    return {k: rel_err(g[k], r[k]) for k in g}


def _problem(seed=7):
    cos = synth.random_draws(1, seed=seed, ragged=False)[0][0]
    p = synth.sixpt_problem(cos)
    rng = np.random.default_rng(seed)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.lens_dz = rng.normal(0, 0.005, p.nl)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    p.bmag = np.linspace(-0.5, 1.0, p.nl)
    return p


@pytest.fixture(scope="module")
def six():
    p = _problem()
    o = O.Oracle().load(p)
    ells = synth.ells_fourier(p.kk_nells, p.kk_lmin, p.kk_lmax)
    return p, o, ells, o.datavector_6x2pt(ells)


@pytest.mark.parametrize("mode", [LEGENDRE_FOLDED, LEGENDRE_FULL])
def test_6x2pt_datavector(six, mode):
    p, o, ells, ref = six
    c = Cosmolike().load(p, mode)
    assert c.datavector_size() == len(ref)
    got = c.compute_6x2pt_real()
    e = dv_err(got, ref, p, len(ells))
    print(mode, e)
    assert max(e.values()) < RTOL
    # the 3-entry entry point gives the same vector (gk, ks, kk after the 3x2pt blocks)
    assert np.array_equal(c.compute_3x2pt_real(), got)


def test_6x2pt_block_order(six):
    p, o, ells, ref = six
    ord6 = (3, 4, 5, 0, 1, 2)  # involutive: the reference formula tiles only those
    c = Cosmolike().load(p)
    got = c.compute_6x2pt_real(ord6)
    want = o.datavector_6x2pt(ells, ord6)
    assert rel_err(got.reshape(1, -1), want.reshape(1, -1)) < RTOL
    with pytest.raises(CosmolikeError):
        c.compute_6x2pt_real((1, 2, 0, 3, 4, 5))  # a 3-cycle: blocks would not tile


def _forward(c, n, B):
    import torch
    out = torch.full((B * n,), float("nan"), dtype=torch.float64, device="cuda:0")
    c.batch_forward(out.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(B, n)


def test_6x2pt_batch_and_graph(six):
    p, o, ells, ref = six
    draws = [_problem(seed=s) for s in (31, 32, 33)]
    c = Cosmolike().load(draws[0])
    c.batch_upload(draws)
    n = c.datavector_size()
    got = _forward(c, n, len(draws))
    for i, q in enumerate(draws):
        e = dv_err(got[i], O.Oracle().load(q).datavector_6x2pt(ells), q, len(ells))
        print(i, e)
        assert max(e.values()) < RTOL, i
    c.batch_use_graph(True)
    for _ in range(2):
        assert np.array_equal(_forward(c, n, len(draws)), got)
    c.batch_use_graph(False)


def test_6x2pt_fp32(six):
    p, o, ells, ref = six
    c = Cosmolike().load(p)
    c.set_precision("fp32")
    c.batch_upload([p])
    e = dv_err(_forward(c, c.datavector_size(), 1)[0], ref, p, len(ells))
    print(e)
    assert max(e.values()) < 1e-4


@pytest.mark.parametrize("binning", ["smooth", "onehot"])
def test_6x2pt_kk_bandpower(six, binning):
    """init_cmb_auto_bandpower: the kk block as band powers
    sum_L M[j][L - lmin] C_kk(L) - offset[j]; the five other blocks unchanged.
    The one-hot binning puts single L on both sides of LMIN_tab = 20 (direct
    rows below, table above); the smooth one spans L = 8..400."""
    p, o, ells, ref = six
    if binning == "smooth":
        bp = synth.kk_bandpower_binning()
    else:
        lmin, lmax = 5, 40
        picks = [5, 12, 19, 20, 21, 33, 40]
        M = np.zeros((len(picks), lmax - lmin + 1))
        for j, L in enumerate(picks):
            M[j, L - lmin] = 1.0
        bp = (len(picks), lmin, lmax, M.reshape(-1), np.linspace(-2e-9, 2e-9, len(picks)))
    q = _problem()
    q.kk_bp = bp
    oq = O.Oracle().load(q)
    want = oq.datavector_6x2pt(ells)
    c = Cosmolike().load(q)
    assert c.kk_bandpower_info()[:3] == bp[:3]
    assert c.datavector_size() == len(want) == len(ref) - len(ells) + bp[0]
    got = c.compute_6x2pt_real()
    n5 = len(ref) - len(ells)
    assert rel_err(got[:n5].reshape(1, -1), ref[:n5].reshape(1, -1)) < RTOL
    e = np.abs(got[n5:] - want[n5:]) / np.abs(want[n5:])
    print(binning, e.max())
    assert e.max() < RTOL
    # batched with graph replay, and the block order moves kk whole
    c.batch_upload([q, q])
    n = c.datavector_size()
    b = _forward(c, n, 2)
    assert rel_err(b, np.stack([want, want])) < RTOL
    c.batch_use_graph(True)
    assert np.array_equal(_forward(c, n, 2), b)
    c.batch_use_graph(False)
    rev = c.compute_6x2pt_real((3, 4, 5, 0, 1, 2))
    assert rel_err(rev.reshape(1, -1), oq.datavector_6x2pt(ells, (3, 4, 5, 0, 1, 2)).reshape(1, -1)) < RTOL


def test_6x2pt_back_to_3x2pt(six):
    p, o, ells, ref = six
    c = Cosmolike().load(p)
    assert c.datavector_size() == len(ref)
    c.init_probes("3x2pt")
    assert c.datavector_size() == len(ref) - (p.nl + p.ns) * p.ntheta - len(ells)


def test_6x2pt_mask_with_bandpowers():
    """IP::set_mask on the 6x2pt vector (generic_interface.hpp:895-921): masked
    entries of every block, band powers included, are 0; the others are the
    unmasked values."""
    q = _problem(seed=41)
    q.kk_bp = synth.kk_bandpower_binning()
    ells = synth.ells_fourier(q.kk_nells, q.kk_lmin, q.kk_lmax)
    want = O.Oracle().load(q).datavector_6x2pt(ells)
    c = Cosmolike().load(q)
    n = c.datavector_size()
    assert n == len(want)
    mask = (np.random.default_rng(41).random(n) > 0.3).astype(np.int32)
    mask[-3:] = (0, 1, 0)  # band powers on both sides
    c.like_set_mask(mask)
    got = c.compute_6x2pt_real()
    assert np.all(got[mask == 0] == 0.0)
    keep = mask == 1
    assert rel_err(got[keep].reshape(1, -1), want[keep].reshape(1, -1)) < RTOL
    c.like_set_mask(None)
    assert rel_err(c.compute_6x2pt_real().reshape(1, -1), want.reshape(1, -1)) < RTOL


def test_6x2pt_ragged_layouts():
    """Draws of different 6x2pt layouts in one batch (per-draw captured
    layouts): C_kk at like.ell with 26 theta bins, and band powers with 20
    theta bins; each draw against the oracle with its own layout."""
    qa = _problem(seed=51)
    qb = _problem(seed=52)
    qb.ntheta = 20
    qb.kk_bp = synth.kk_bandpower_binning(nbp=6, lmin=10, lmax=300)
    refs, nfs, lays = [], [], []
    c = Cosmolike()
    for q in (qa, qb):
        ells = synth.ells_fourier(q.kk_nells, q.kk_lmin, q.kk_lmax)
        refs.append(O.Oracle().load(q).datavector_6x2pt(ells))
        nfs.append(q.kk_bp[0] if q.kk_bp is not None else len(ells))
        c.load(q)
        lays.append(c.capture_layout())
    order = [0, 1, 0, 1]
    c.batch_upload_layouts([(qa, qb)[k] for k in order], [lays[k] for k in order])
    offs = c.batch_offsets()
    assert offs[-1] == sum(len(refs[k]) for k in order)
    got = _forward(c, offs[-1], 1)[0]
    for i, k in enumerate(order):
        e = dv_err(got[offs[i]:offs[i + 1]], refs[k], (qa, qb)[k], nfs[k])
        print(i, k, e)
        assert max(e.values()) < RTOL, (i, k)


def test_6x2pt_chi2():
    """IP::get_chi2 on the 6x2pt vector (band powers and a mask included),
    host entry and batched device entry with its gradient, against numpy."""
    import torch
    q = _problem(seed=81)
    q.kk_bp = synth.kk_bandpower_binning()
    c = Cosmolike().load(q)
    n = c.datavector_size()
    rng = np.random.default_rng(81)
    mask = (rng.random(n) > 0.2).astype(np.int32)
    c.like_set_mask(mask)
    theory = c.compute_6x2pt_real()
    sig = 0.05 * np.abs(theory) + 1e-12
    data = theory + sig * rng.standard_normal(n)
    icov = np.zeros((n, n))
    icov[np.arange(n), np.arange(n)] = mask / sig**2
    c.like_set_data(data)
    c.like_set_inv_cov(icov)
    r = (data - theory) * mask
    ref = float(r @ (icov @ r))
    assert abs(c.chi2(theory) / ref - 1) < 1e-12
    c.batch_upload([q, q])
    dv = torch.tensor(np.concatenate([theory, data]), dtype=torch.float64, device="cuda:0")
    chi = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    grad = torch.zeros(2 * n, dtype=torch.float64, device="cuda:0")
    c.batch_chi2(dv.data_ptr(), chi.data_ptr(), grad.data_ptr())
    torch.cuda.synchronize()
    got = chi.cpu().numpy()
    assert abs(got[0] / ref - 1) < 1e-12 and abs(got[1]) < 1e-20 * max(ref, 1.0)
    g = grad.cpu().numpy()[:n]
    assert np.allclose(g, -2 * icov @ r, rtol=1e-12, atol=1e-30)
