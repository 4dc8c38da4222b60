"""cfd_sip9p_noalloc (the p' solve that allocates nothing, its buffers in the library's
static device array) against cfd_sip9p through the C ABI: two calls in a row, the second
on a smaller system, each equal to cfd_sip9p bit for bit (same kernels, same order)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(name, A9, phi0, q, rtol, maxit):
    from cfd_solver_amd._lib import call, f64
    from cfd_solver_amd.stages import K9, SIP_ORDER
    names9 = {"p": (0, 0), "e": (1, 0), "w": (-1, 0), "n": (0, 1), "s": (0, -1), "ne": (1, 1),
              "nw": (-1, 1), "se": (1, -1), "sw": (-1, -1)}
    planes = {nm: f64(A9[K9[o]]) for nm, o in names9.items()}
    n1, n0 = q.shape
    phi = f64(phi0).copy()
    it, rr = C.c_int(), C.c_double()
    call(name, n0, n1, *[planes[nm] for nm in SIP_ORDER], phi, f64(q), rtol, maxit, C.byref(it), C.byref(rr))
    return phi, it.value, rr.value


def _ogrid_system(n0, n1, seed):
    """The reference's 9-point p' operator (solver.cpp:943-1039, oracle restatement) on an
    O-grid around an ellipse, with a seeded right-hand side on the unknown rows."""
    from cfd_solver_amd.grid import ogrid_ellipse
    from oracle.oracle import lib
    g = ogrid_ellipse(n0, n1, ar=2.0)
    A9 = np.zeros((9, n1, n0))
    c = np.ascontiguousarray
    lib().or_assemble_p9_ref(n0, n1, g.dxi, g.deta, c(g.xix), c(g.xiy), c(g.etx), c(g.ety), A9)
    q = np.zeros((n1, n0))
    q[1:-1, :-1] = np.random.default_rng(seed).standard_normal((n1 - 2, n0 - 1))
    q[1:-1, -1] = q[1:-1, 0]
    return A9, q


def test_noalloc_sip9p_equals_sip9p(golden):
    print("\n[gpu] cfd_sip9p_noalloc vs cfd_sip9p: 129x65 O-grid, then the reference's 33x17 system", flush=True)
    cases = [_ogrid_system(129, 65, 3) + (1e-10,), (golden["A9"], golden["s2_q"], 1e-12)]
    for A9, q, rtol in cases:
        phi0 = np.zeros_like(q)
        a, ia, ra = _solve("cfd_sip9p", A9, phi0, q, rtol, 500)
        b, ib, rb = _solve("cfd_sip9p_noalloc", A9, phi0, q, rtol, 500)
        assert ia == ib and ra == rb and ra <= rtol, (ia, ib, ra, rb)
        assert np.array_equal(a, b)


def test_noalloc_sip9p_rejects_oversize():
    """More nodes than the reference's 350 x 570 arrays: an argument error, nothing run."""
    from cfd_solver_amd._lib import load
    n0, n1 = 351, 571
    z = np.zeros((n1, n0))
    it, rr = C.c_int(), C.c_double()
    code = load().cfd_sip9p_noalloc(n0, n1, *([z] * 11), 1e-8, 10, C.byref(it), C.byref(rr))
    assert code == -1
