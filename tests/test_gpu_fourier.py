"""Fourier-space 3x2pt data vector (compute_Mx2pt_N_masked<1,3>) on the GPU
vs the FP64 oracle: C_ss EE (cut at lmax_shear), C_gs, C_gg at the
init_binning_fourier ells by the direct GL integrals, shear calibration, block
order and mask; also under TATT intrinsic alignments and the one-loop bias.

Bar: 1e-6 relative per entry, with a floor of 1e-3 of each block row's
largest magnitude (metric of the FP64 parity tests); masked and cut entries
exactly 0.
"""
import numpy as np
import pytest

from cosmolike_amd import Cosmolike, synth

import oracle as O

pytestmark = pytest.mark.gpu


def rel_err(gpu, ref, ncl):
    g, r = gpu.reshape(-1, ncl), ref.reshape(-1, ncl)
    floor = 1e-3 * np.abs(r).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(r), floor)
    den[den == 0] = 1.0
    return float((np.abs(g - r) / den).max())


def ells_of(nells, lmin, lmax):
    e = lmin * (lmax / lmin) ** (np.arange(nells + 1) / nells)
    return np.sqrt(e[:-1] * e[1:])


@pytest.mark.parametrize("nells,lmin,lmax,lmax_shear,ord", [
    (15, 30, 3000, 1000, (0, 1, 2)),
    (20, 20, 8000, 3000, (2, 1, 0)),
    (1, 100, 200, 10**6, (0, 2, 1)),
])
def test_fourier_3x2pt_matches_oracle(nells, lmin, lmax, lmax_shear, ord):
    cos = synth.random_draws(1, seed=nells, ragged=False)[0][0]
    p = synth.full_3x2pt_problem(cos)
    rng = np.random.default_rng(nells)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    c = Cosmolike().load(p)
    c.init_binning_fourier(nells, lmin, lmax, lmax_shear)
    ref = O.Oracle().load(p).datavector_fourier(ells_of(nells, lmin, lmax), lmax_shear, ord)
    got = c.compute_3x2pt_fourier(ord)
    assert got.shape == ref.shape
    assert np.array_equal(got == 0, ref == 0)
    e = rel_err(got, ref, nells)
    print(nells, lmin, lmax, lmax_shear, ord, "err", e)
    assert e < 1e-6
    mask = (rng.random(got.size) > 0.3).astype(np.int32)
    gm = c.compute_3x2pt_fourier(ord, mask)
    assert np.all(gm[mask == 0] == 0)
    assert np.array_equal(gm[mask == 1], got[mask == 1])
    # the real-space data vector is unaffected by the Fourier call
    assert rel_err(c.compute_3x2pt_real(), O.Oracle().load(p).datavector(), 26) < 1e-6


@pytest.mark.parametrize("model", ["tatt", "oneloop"])
def test_fourier_3x2pt_models(model):
    """The Fourier-space vector under TATT intrinsic alignments (C_ss EE and
    the TATT C_gs) and under the one-loop galaxy bias (C_gs, C_gg)."""
    cos = synth.random_draws(1, seed=71, ragged=False)[0][0]
    p = synth.tatt_problem(cos) if model == "tatt" else synth.oneloop_problem(cos)
    rng = np.random.default_rng(71)
    p.src_dz = rng.normal(0, 0.01, p.ns)
    p.shear_m = rng.normal(0, 0.01, p.ns)
    nells, lmin, lmax, lmax_shear = 12, 30, 3000, 2000
    c = Cosmolike().load(p)
    c.init_binning_fourier(nells, lmin, lmax, lmax_shear)
    o = O.Oracle().load(p)
    ref = o.datavector_fourier(ells_of(nells, lmin, lmax), lmax_shear, (0, 1, 2))
    got = c.compute_3x2pt_fourier((0, 1, 2))
    e = rel_err(got, ref, nells)
    print(model, "err", e)
    assert e < 1e-6
    # the model terms are really exercised: the NLA / linear-bias vector differs
    q = synth.full_3x2pt_problem(cos)
    q.src_dz, q.shear_m = p.src_dz, p.shear_m
    base = O.Oracle().load(q).datavector_fourier(ells_of(nells, lmin, lmax), lmax_shear, (0, 1, 2))
    assert rel_err(base, ref, nells) > 1e-3
