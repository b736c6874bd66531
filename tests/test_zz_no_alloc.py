"""The allocation-free entry point clh_compute_3x2pt_real_noalloc against
clh_compute_3x2pt_real, through the C ABI.

Two calls in a row on one context, the second with a smaller layout (fewer
bins, fewer theta bins) than the first, so the second finds the first's
contents in the shared scratch array; each matches the allocating entry point
on the same inputs.  Also: the call is refused before the layout's device copy
exists (CLH_ERR_STATE) and for a draw that does not fit the 16 MiB array
(CLH_ERR_ARG), leaving the output untouched in both cases.

Bar: 1e-12 relative to the allocating entry point (the same kernels on the
same inputs; the FP64 parity bar against the oracle is 1e-6).
"""
import ctypes

import numpy as np
import pytest
from scipy import interpolate

from cosmolike_amd import Cosmolike, synth

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_STATE = -1, -3


def _noalloc(c, ord=(0, 1, 2), n=None, out=None):
    fn = c.L.clh_compute_3x2pt_real_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    n = c.datavector_size() if n is None else n
    out = np.empty(n) if out is None else out
    o = (ctypes.c_int * 3)(*ord)
    rc = fn(c.h, o, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return rc, out


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * np.abs(b).max())))


def test_noalloc_matches_allocating_entry_twice_then_smaller():
    draws = synth.random_draws(2, seed=77, ragged=False)
    big = synth.full_3x2pt_problem(draws[0][0])
    rng = np.random.default_rng(77)
    big.src_dz = rng.normal(0, 0.01, big.ns)
    big.shear_m = rng.normal(0, 0.01, big.ns)
    small = synth.ragged_3x2pt_problem(draws[1][0], 4, 3, 20)

    c = Cosmolike(device=0)
    c.load(big)
    # before any compute the layout has no device copy: refused, output untouched
    sentinel = np.full(c.datavector_size(), 5.0)
    rc, _ = _noalloc(c, out=sentinel)
    assert rc == ERR_STATE and np.all(sentinel == 5.0)

    for p, ord in ((big, (0, 1, 2)), (small, (2, 1, 0))):
        c.load(p)
        ref = c.compute_3x2pt_real(ord)
        rc, got = _noalloc(c, ord)
        assert rc == 0, c.L.clh_last_error(c.h)
        e = _rel(got, ref)
        print(p.ns, p.nl, p.ntheta, ord, "noalloc vs allocating", e, "bit-identical", np.array_equal(got, ref))
        assert e < 1e-12
        # a second call on the same inputs gives the same vector
        rc, again = _noalloc(c, ord)
        assert rc == 0 and np.array_equal(again, got)


def test_noalloc_refuses_a_draw_beyond_the_scratch():
    p = synth.full_3x2pt_problem(synth.random_draws(1, seed=78, ragged=False)[0][0])
    c = Cosmolike(device=0).load(p)
    ref = c.compute_3x2pt_real()
    # the same layout with a 4000 x 600 ln P(k,z) table (19 MB > 16 MiB)
    lk = np.linspace(p.log10k[0], p.log10k[-1], 4000)
    zz = np.linspace(p.pk_z[0], p.pk_z[-1], 600)
    rgi = interpolate.RegularGridInterpolator((p.log10k, p.pk_z), p.lnP.reshape(len(p.log10k), len(p.pk_z)))
    L, Z = np.meshgrid(lk, zz, indexing="ij")
    c.set_non_linear_power_spectrum(lk, zz, rgi(np.c_[L.ravel(), Z.ravel()]))
    sentinel = np.full(c.datavector_size(), 5.0)
    rc, _ = _noalloc(c, out=sentinel)
    assert rc == ERR_ARG and np.all(sentinel == 5.0)
    assert b"scratch" in c.L.clh_last_error(c.h)
    # back to the draw's own table: the entry point works again
    c.set_non_linear_power_spectrum(p.log10k, p.pk_z, p.lnP)
    rc, got = _noalloc(c)
    assert rc == 0 and _rel(got, ref) < 1e-12
