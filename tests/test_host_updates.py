"""Host-side optimizer of the reference (sjcgsolver, main.cpp:1232-1303) in
libds2d.so, checked against hand-computed values (no device work)."""
import numpy as np

from ds2d_amd.lib import sjcg_update


def test_sjcg_first_iteration_is_scaled_steepest_descent():
    m = np.array([1.0, 2.0], np.float32)
    cg = np.array([7.0, -3.0], np.float32)   # ignored at iteration 0 (beta = 0)
    g1 = np.array([0.5, -1.0], np.float32)
    g0 = np.zeros(2, np.float32)
    a, b = sjcg_update(m, cg, g1, g0, 0)
    assert b == 0.0 and a == np.float32(1e-4)
    np.testing.assert_array_equal(cg, -g1)
    np.testing.assert_array_equal(g0, g1)
    np.testing.assert_allclose(m, [1.0 - 0.5e-4, 2.0 + 1e-4], rtol=1e-7)


def test_sjcg_hybrid_beta_and_step_length():
    m = np.array([1.0, 2.0], np.float32)
    cg = np.array([-0.5, 1.0], np.float32)
    g0 = np.array([0.5, -1.0], np.float32)
    g1 = np.array([1.0, 1.0], np.float32)
    m_before = m.copy()
    a, b = sjcg_update(m, cg, g1, g0, 1)
    # g1 - g0 = [0.5, 2]: HS = 2.5 / 1.75, DY = 2 / 1.75 -> beta = min = 8/7
    assert abs(b - 8.0 / 7.0) < 1e-6
    np.testing.assert_allclose(cg, [-1.0 - 0.5 * 8 / 7, -1.0 + 8 / 7], rtol=1e-6)
    # alpha = 0.1 * max|m| / max|cg| (model before the update)
    assert abs(a - 0.1 * 2.0 / (1.0 + 4.0 / 7.0)) < 1e-6
    np.testing.assert_allclose(m, m_before + a * cg, rtol=1e-6)


def test_sjcg_beta_is_clipped_at_zero():
    m = np.array([1.0, -3.0], np.float32)
    cg = np.array([1.0, 0.0], np.float32)
    g0 = np.array([0.0, 0.0], np.float32)
    g1 = np.array([-1.0, 0.0], np.float32)  # b = cg.(g1-g0) = -1 -> HS = 1/-1 < 0 -> beta 0
    _, b = sjcg_update(m, cg, g1, g0, 3)
    assert b == 0.0
    np.testing.assert_array_equal(cg, [1.0, 0.0])
