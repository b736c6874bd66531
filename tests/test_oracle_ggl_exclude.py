"""gs pair exclusion in the oracle (init_ggl_exclude,
generic_interface.cpp:874-896; test_zoverlap / N_ggl, redshift_spline.c:115-223):
the data vector with excluded (lens, source) pairs equals the one without
exclusions with those pairs' gamma_t rows taken out, the rest in lens-major
order (ZL / ZS of the kept pairs), in real and Fourier space, for block
orders that move the gamma_t block, with the point mass on."""
import numpy as np
import pytest

from cosmolike_amd import synth

import oracle as O

EXCL = [(0, 0), (0, 3), (2, 9), (4, 1), (4, 9)]


def _drop(dv, nss_rows, ns, nl, nrow_len, ord):
    """Blocks of dv in the order ord -> the same vector with the excluded gs rows removed."""
    sizes = [nss_rows * nrow_len, ns * nl * nrow_len, nl * nrow_len]
    idx = np.argsort(ord, kind="stable")
    start = [sum(sizes[idx[j]] for j in range(list(idx).index(i))) for i in range(3)]
    blocks = [dv[start[i]:start[i] + sizes[i]] for i in range(3)]
    keep = [zl * ns + zs for zl in range(nl) for zs in range(ns) if (zl, zs) not in EXCL]
    blocks[1] = blocks[1].reshape(ns * nl, nrow_len)[keep].ravel()
    return np.concatenate([blocks[i] for i in idx])


@pytest.fixture(scope="module")
def prob():
    p = synth.full_3x2pt_problem()
    p.pm = np.linspace(0.5, 2.0, p.nl)
    p.shear_m = np.linspace(-0.02, 0.03, p.ns)
    return p


@pytest.mark.parametrize("ord", [(0, 1, 2), (2, 1, 0), (1, 0, 2)])
def test_real_space_exclusion_drops_rows(prob, ord):
    full = O.Oracle().load(prob).datavector(ord)
    o = O.Oracle().load(prob)
    o.set_ggl_exclude(np.array(EXCL).ravel())
    assert o.n_ggl() == 45
    assert all(o.test_zoverlap(zl, zs) == ((zl, zs) not in EXCL) for zl in range(5) for zs in range(10))
    got = o.datavector(ord)
    assert got.size == full.size - 5 * prob.ntheta
    assert np.array_equal(got, _drop(full, 2 * 55, 10, 5, prob.ntheta, ord))


def test_fourier_exclusion_drops_rows(prob):
    ells = np.geomspace(30, 3000, 6)
    for ord in ((0, 1, 2), (0, 2, 1)):
        full = O.Oracle().load(prob).datavector_fourier(ells, 1000.0, ord)
        o = O.Oracle().load(prob)
        o.set_ggl_exclude(np.array(EXCL).ravel())
        got = o.datavector_fourier(ells, 1000.0, ord)
        assert np.array_equal(got, _drop(full, 55, 10, 5, len(ells), ord))
