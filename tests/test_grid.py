"""Grid generator, metrics and the INP.DAT format."""
import math
import os

import numpy as np

from cfd_solver_amd.grid import cavity_grid, ogrid_ellipse, read_inp_dat, round_pow2p1, write_inp_dat


def test_skew_jacobian_closed_form():
    # x = xi + A s, y = eta + A s with s = sin(pi xi) sin(pi eta): J = 1 + A pi sin(pi (xi + eta))
    n, A = 129, 0.1
    g = cavity_grid(n, skew=A, beta=None, control=False)
    xi = np.linspace(0, 1, n)
    XI, ET = np.meshgrid(xi, xi)
    exact = 1.0 + A * math.pi * np.sin(math.pi * (XI + ET))
    err = np.abs(g.ajac - exact)[1:-1, 1:-1].max()
    assert err < 2e-3  # second-order differences, h = 1/128
    g2 = cavity_grid(2 * n - 1, skew=A, beta=None, control=False)
    XI2, ET2 = np.meshgrid(np.linspace(0, 1, 2 * n - 1), np.linspace(0, 1, 2 * n - 1))
    err2 = np.abs(g2.ajac - (1.0 + A * math.pi * np.sin(math.pi * (XI2 + ET2))))[1:-1, 1:-1].max()
    assert 3.5 < err / err2 < 4.5  # O(h^2)


def test_metric_identities():
    g = cavity_grid(65, skew=0.15, beta=1.5)
    np.testing.assert_allclose(g.alph, g.xix ** 2 + g.xiy ** 2, rtol=1e-14)
    np.testing.assert_allclose(g.gamma, g.etx ** 2 + g.ety ** 2, rtol=1e-14)
    np.testing.assert_allclose(g.beta, -(g.xix * g.etx + g.xiy * g.ety), rtol=1e-12, atol=1e-14)
    assert g.ajac.min() > 0
    # straight walls: eta_x = 0 along y = 0 and y = 1, xi_y = 0 along x = 0 and x = 1
    assert np.abs(g.etx[0]).max() < 1e-12 and np.abs(g.etx[-1]).max() < 1e-12
    assert np.abs(g.xiy[:, 0]).max() < 1e-12 and np.abs(g.xiy[:, -1]).max() < 1e-12


def test_control_functions_cartesian_zero_and_stretched():
    g = cavity_grid(33, skew=0.0, beta=None)
    assert np.abs(g.p1).max() < 1e-10 and np.abs(g.q1).max() < 1e-10
    # 1-D stretch: P = lap(xi) = -x''(xi)/x'(xi)^3 along xi
    g = cavity_grid(257, skew=0.0, beta=1.5)
    b = 1.5
    xi = np.linspace(0, 1, 257)
    xp = b / math.tanh(b) / np.cosh(b * (2 * xi - 1)) ** 2
    xpp = -4 * b * b / math.tanh(b) * np.tanh(b * (2 * xi - 1)) / np.cosh(b * (2 * xi - 1)) ** 2
    exact = -xpp / xp ** 3
    np.testing.assert_allclose(g.p1[128, 1:-1], exact[1:-1], rtol=2e-3, atol=1e-3)


def test_ogrid_periodic_duplicate():
    g = ogrid_ellipse(33, 17, ar=2.0)
    for a in (g.x, g.y, g.xix, g.xiy, g.etx, g.ety, g.alph, g.beta, g.gamma, g.ajac):
        np.testing.assert_array_equal(a[:, -1], a[:, 0])
    assert g.ajac.min() > 0


def test_inp_dat_round_trip(tmp_path):
    g = ogrid_ellipse(17, 9, ar=1.5)
    p = os.path.join(tmp_path, "INP.DAT")
    write_inp_dat(p, g)
    h = read_inp_dat(p)
    assert (h.n0, h.n1) == (17, 9)
    for k in ("x", "y", "xix", "xiy", "etx", "ety", "alph", "beta", "gamma", "ajac"):
        np.testing.assert_array_equal(getattr(h, k), getattr(g, k))
    np.testing.assert_array_equal(h.xnox, g.xnox)
    assert h.dxi == g.dxi and h.ar == g.ar
    assert not h.p1.any() and not h.q1.any()


def test_inp_dat_truncated(tmp_path):
    g = ogrid_ellipse(9, 5)
    p = os.path.join(tmp_path, "INP.DAT")
    write_inp_dat(p, g)
    txt = open(p).read().split("\n")
    open(p, "w").write("\n".join(txt[: len(txt) // 2]))
    import pytest
    with pytest.raises(ValueError):
        read_inp_dat(p)


def test_round_pow2p1():
    assert [round_pow2p1(n) for n in (65, 100, 129, 200, 513, 1025)] == [65, 129, 129, 257, 513, 1025]
