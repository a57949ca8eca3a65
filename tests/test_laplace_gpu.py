"""GPU Laplacian filters (ds2d_laplace_filter / ds2d_laplace_band) against the
reference's own laplace.cpp (compiled into oracle/_ref from its source) and
the oracle.  Integer-free float filters with the reference's arithmetic order:
bit-identical results are expected."""
import numpy as np
import pytest

import oracle_lib as orc
from ds2d_amd import laplace_band, laplace_filter

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,dx", [((40, 23), 10.0), ((251, 101), 10.0), ((512, 256), 12.5), ((5, 5), 1.0)])
def test_laplace_filter_bit_identical_to_reference(shape, dx):
    rng = np.random.default_rng(shape[0])
    img = rng.standard_normal(shape).astype(np.float32)
    out = laplace_filter(img, dx)
    ours = orc.laplace_order(img, dx, prec="f")
    assert np.array_equal(out, ours)
    if orc.ref_laplace_available():
        assert np.array_equal(out, orc.ref_laplace_order(img, dx))
    # closed form on ix^2 + iz^2: 4 / dx^2 inside, zero frame
    ix, iz = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    q = laplace_filter((ix ** 2 + iz ** 2).astype(np.float32), dx)
    if shape[0] > 4 and shape[1] > 4:
        np.testing.assert_allclose(q[2:-2, 2:-2], 4 / dx ** 2, rtol=1e-6)
    fr = np.ones(shape, bool)
    fr[2:-2, 2:-2] = False
    assert np.all(q[fr] == 0)


@pytest.mark.parametrize("shape", [(40, 36), (251, 101), (512, 256), (3, 2)])
def test_laplace_band_matches_oracle(shape):
    rng = np.random.default_rng(shape[1])
    img = rng.standard_normal(shape).astype(np.float32)
    out = laplace_band(img)
    assert np.array_equal(out, orc.laplace_band(img, prec="f"))
    ref = orc.laplace_band(img.astype(np.float64))
    assert orc.rel_l2(out, ref) <= 1e-6
