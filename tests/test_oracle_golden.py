"""The oracle's reference-mode restatement against the compiled reference's own arrays.

Each stage is fed the reference's inputs from the fixture and must reproduce
the reference's outputs to rounding (in practice bit for bit).
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import lib

INT = np.s_[1:-1, :]  # rows the reference writes


def _g(d):
    return {k: d["grid_" + k] for k in ["xix", "xiy", "etx", "ety", "alph", "beta", "gamma", "ajac", "x", "y"]}


def test_assemble_p9(golden):
    d = golden
    g = _g(d)
    A9 = np.zeros((9, d["n1"], d["n0"]))
    lib().or_assemble_p9_ref(d["n0"], d["n1"], d["dxi"], d["deta"], g["xix"], g["xiy"], g["etx"], g["ety"], A9)
    assert np.array_equal(A9[:, 1:-1, :], d["A9"][:, 1:-1, :])


def test_assemble_mom(golden):
    d = golden
    g = _g(d)
    n0, n1 = d["n0"], d["n1"]
    AU, AT = np.zeros((25, n1, n0)), np.zeros((25, n1, n0))
    BU, BT = np.zeros((6, n0)), np.zeros((6, n0))
    z = np.zeros((n1, n0))
    lib().or_assemble_mom_ref(n0, n1, d["dxi"], d["deta"], g["alph"], g["beta"], g["gamma"], z, z,
                              d["dt"], d["Re"], d["Pr"], AU, AT, BU, BT)
    assert np.array_equal(AU[:, 1:-1, :], d["AU"][:, 1:-1, :])
    assert np.array_equal(AT[:, 1:-1, :], d["AT"][:, 1:-1, :])
    assert np.array_equal(BU[:, :n0 - 1], d["BU"][:, :n0 - 1])
    assert np.array_equal(BT[:, :n0 - 1], d["BT"][:, :n0 - 1])


def test_rhs_step2(golden):
    d = golden
    g = _g(d)
    n0, n1 = d["n0"], d["n1"]
    out = [np.zeros((n1, n0)) for _ in range(7)]
    lib().or_rhs_ref(n0, n1, d["dxi"], d["deta"], d["s2_uold"], d["s2_vold"], d["s2_Told"], d["s1_up"],
                     d["s1_vp"], d["s1_p"], g["xix"], g["xiy"], g["etx"], g["ety"], g["alph"], g["gamma"],
                     d["BU"], d["BT"], d["Re"], d["Pr"], float(d["s2_Ri"][0]), d["dt"], *out)
    for k, o in zip(["uxi", "uet", "qu", "qv", "qt", "qup", "qvp"], out):
        if k in ("uxi", "uet"):
            continue  # the reference recomputes these later in the step
        assert np.array_equal(o[INT], d["s2_" + k][INT]), k


def test_gauss_single_call(golden):
    d = golden
    n0, n1 = d["n0"], d["n1"]
    r = C.c_double()
    for name, A, q, start, quirk, ref in [("u", d["AU"], d["s2_qu"], d["s2_uold"], 0, d["s2_us"]),
                                          ("v", d["AU"], d["s2_qv"], d["s2_vold"], 0, d["s2_vs"]),
                                          ("T", d["AT"], d["s2_qt"], d["s2_Told"], 1, d["s2_T"])]:
        phi = start.copy()
        lib().or_gauss_ref(n0, n1, np.ascontiguousarray(A), quirk, phi, q, 0.75e-2, 100000, C.byref(r))
        assert np.array_equal(phi[INT], ref[INT]), name


def test_face_div(golden):
    d = golden
    g = _g(d)
    q = np.zeros((d["n1"], d["n0"]))
    lib().or_face_div_ref(d["n0"], d["n1"], d["dxi"], d["deta"], d["s2_up"], d["s2_vp"], d["s1_p"],
                          g["xix"], g["xiy"], g["etx"], g["ety"], d["dt"], q)
    assert np.array_equal(q[1:-1, :-1], d["s2_q"][1:-1, :-1])


def test_sip9p_single_call_and_converged(golden):
    d = golden
    n0, n1 = d["n0"], d["n1"]
    r = C.c_double()
    phi = np.zeros((n1, n0))
    lib().or_sip9p_ref(n0, n1, d["A9"], phi, d["s2_q"], 0.75e-2, 100000, 0.92, C.byref(r))
    assert np.array_equal(phi, d["s2_pcor_one"])
    assert np.array_equal(phi[1:-1, :], d["s2_pcor"][1:-1, :])
    phi = np.zeros((n1, n0))
    for _ in range(7):
        lib().or_sip9p_ref(n0, n1, d["A9"], phi, d["s2_q"], 0.75e-2, 100000, 0.92, C.byref(r))
    assert np.array_equal(phi, d["s2_pcor_conv"])


def test_pcor_bc_and_correct(golden):
    d = golden
    g = _g(d)
    n0, n1 = d["n0"], d["n1"]
    pc = d["s2_pcor_one"].copy()
    lib().or_pcor_bc_ref(n0, n1, pc, d["grid_xnox"], d["grid_xnoy"], d["uinf"], d["vinf"])
    assert np.array_equal(pc, d["s2_pcor"])
    u, v, p = np.zeros((n1, n0)), np.zeros((n1, n0)), d["s1_p"].copy()
    lib().or_correct_ref(n0, n1, d["dxi"], d["deta"], d["s2_us"], d["s2_vs"], d["s2_pcor"], g["xix"],
                         g["xiy"], g["etx"], g["ety"], d["dt"], u, v, p)
    assert np.array_equal(u[INT], d["s2_u"][INT])
    assert np.array_equal(v[INT], d["s2_v"][INT])
    assert np.array_equal(p[INT], d["s2_p"][INT])


def test_wall_pressure(golden):
    d = golden
    g = _g(d)
    n0, n1 = d["n0"], d["n1"]
    p = d["s2_p"].copy()
    p[0, :] = 0.0
    z = np.zeros((n1, n0))
    z0 = np.zeros(n0)
    lib().or_wall_p_ref(n0, n1, d["dxi"], d["deta"], d["s2_u"], d["s2_v"], d["s2_T"], p, g["xix"], g["xiy"],
                        g["etx"], g["ety"], g["alph"], g["beta"], g["gamma"], z, g["ajac"], z0, z0,
                        d["Re"], float(d["s2_Ri"][0]))
    assert np.array_equal(p[0], d["s2_p"][0])


def test_dilation(golden):
    d = golden
    g = _g(d)
    dil = np.zeros((d["n1"], d["n0"]))
    dm = lib().or_dil_ref(d["n0"], d["n1"], d["dxi"], d["deta"], d["s2_u"], d["s2_v"], g["xix"], g["xiy"],
                          g["etx"], g["ety"], dil)
    assert np.array_equal(dil[INT], d["s2_dil"][INT])
    assert dm == d["s2_dmax"][0]


def test_nusselt(golden):
    d = golden
    g = _g(d)
    prof = np.zeros(d["n0"])
    nu = lib().or_nusselt_ref(d["n0"], d["n1"], d["dxi"], d["deta"], d["s0_T"], g["ajac"], g["etx"],
                              g["ety"], d["ar"], prof.ctypes.data)
    # COEFF_HIS.dat prints Nu with 8 decimals, SURF_DIST.dat dT/dn with 6 significant digits
    assert abs(nu - d["s0_nu_8dp"][0]) <= 5e-9
    np.testing.assert_allclose(prof, d["s0_dthdn_6sig"], rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("mistake", ["sign", "index"])
def test_golden_catches_planted_mistake(golden, mistake):
    """A planted oracle-side mistake must break the golden comparison."""
    d = golden
    g = _g(d)
    q = np.zeros((d["n1"], d["n0"]))
    up = d["s2_up"].copy()
    if mistake == "sign":
        vp = -d["s2_vp"]
    else:
        vp = np.roll(d["s2_vp"], 1, axis=1)
    lib().or_face_div_ref(d["n0"], d["n1"], d["dxi"], d["deta"], up, vp, d["s1_p"], g["xix"], g["xiy"],
                          g["etx"], g["ety"], d["dt"], q)
    assert not np.allclose(q[1:-1, :-1], d["s2_q"][1:-1, :-1], rtol=1e-6)


def test_golden_pcor_conv_equals_dense_solve(golden):
    """The reference's repeated-SIP solution is the exact solution of its 9-point system."""
    d = golden
    A9, q, conv = d["A9"], d["s2_q"], d["s2_pcor_conv"]
    n1, n0 = q.shape
    P = n0 - 1

    def k(i, j):
        return (j - 1) * P + i

    M = np.zeros((P * (n1 - 2), P * (n1 - 2)))
    b = np.zeros(P * (n1 - 2))
    for j in range(1, n1 - 1):
        for i in range(P):
            b[k(i, j)] = q[j, i]
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    if 1 <= j + dj <= n1 - 2:
                        M[k(i, j), k((i + di) % P, j + dj)] += A9[(dj + 1) * 3 + di + 1, j, i]
    x = np.linalg.solve(M, b).reshape(n1 - 2, P)
    assert np.linalg.norm(x - conv[1:-1, :-1]) <= 1e-13 * np.linalg.norm(x)


# ---- outer-boundary stages and the whole time step ------------------------------------------

def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def oracle_farfield(a, cos_ft=1.0):
    n1, n0 = a["u"].shape
    o = {k: _c(a[k]).copy() for k in ("u", "v", "T", "up", "vp")}
    circ = lib().or_farfield_ref(n0, n1, a["deta"], _c(a["x"]), _c(a["y"]), o["u"], o["v"], o["T"], o["up"],
                                 o["vp"], _c(a["Told"]), _c(a["uet"]), _c(a["etx"]), _c(a["ety"]), _c(a["ajac"]),
                                 _c(a["xnox"]), _c(a["xnoy"]), a["uinf"], a["vinf"], a["dt"], a["speed_amp"], cos_ft)
    return o, circ


def test_outflow_up(golden):
    from tests.ref_bc_cases import outflow_golden
    a, (wu, wv) = outflow_golden(golden)
    up, vp = a["up"].copy(), a["vp"].copy()
    lib().or_outflow_up_ref(golden["n0"], golden["n1"], up, vp, _c(a["u"]), _c(a["v"]), _c(a["xnox"]),
                            _c(a["xnoy"]), a["uinf"], a["vinf"])
    assert np.array_equal(up, wu) and np.array_equal(vp, wv)


def test_farfield(golden):
    from tests.ref_bc_cases import farfield_golden
    a, want = farfield_golden(golden)
    o, circ = oracle_farfield(a)
    for k, w in want.items():
        assert np.array_equal(o[k], w), k
    assert abs(circ) < 1e-2  # this step takes the k = 2 branch; the closed form covers k = 1


def test_exit_pressure(golden):
    from tests.ref_bc_cases import exit_p_golden
    d, g = golden, _g(golden)
    a, want = exit_p_golden(d)
    p = a["p"].copy()
    z = np.zeros((d["n1"], d["n0"]))
    lib().or_exit_p_ref(d["n0"], d["n1"], d["dxi"], d["deta"], _c(a["u"]), _c(a["v"]), _c(a["T"]), _c(a["uold"]),
                        _c(a["vold"]), p, g["xix"], g["xiy"], g["etx"], g["ety"], g["alph"], g["beta"], g["gamma"],
                        z, g["ajac"], _c(a["xnox"]), _c(a["xnoy"]), a["uinf"], a["vinf"], a["dt"], a["Re"], a["Ri"])
    assert np.array_equal(p, want)
    lv = a["uinf"] * a["xnox"] + a["vinf"] * a["xnoy"]
    assert (lv[:-1] >= 0).any() and (lv[:-1] < 0).any()  # both branches are exercised


def test_full_step_reproduces_reference_step(golden):
    """Every stage composed in the reference's order, from the end of step 1, rebuilds the
    reference's step 2 bit for bit -- interior and boundary rows, including the loose
    Gauss-Seidel and SIP solves."""
    from oracle.oracle import ref_step
    from tests.ref_bc_cases import golden_step_inputs
    o = ref_step(*golden_step_inputs(golden))
    for k in ("u", "v", "T", "p", "up", "vp", "pcor", "q", "dil", "uold", "vold", "Told"):
        assert np.array_equal(o[k], golden["s2_" + k]), k
    assert o["dmax"] == golden["s2_dmax"][0]


@pytest.mark.parametrize("G", [1.0, 1e-3])
def test_farfield_closed_form(G):
    """Free stream + A/r^2 radial + G/r vortex on two polar rows: the k = 1 branch
    (|circulation| > 1e-2) carries the vortex exactly as G/r, the k = 2 branch as
    G r0 / r1^2; the wall row moves as a rigid rotation; at CFL 1 the outflow
    temperature is the previous row's old value."""
    from tests.ref_bc_cases import polar_case
    a, (ou, ov, oT), (wu, wv), leave, k, circ_est = polar_case(G)
    assert k == (1 if G > 0.1 else 2) and leave.any() and (~leave[:-1]).any()
    o, circ = oracle_farfield(a)
    assert abs(circ - circ_est) <= 1e-12 * abs(circ_est)
    np.testing.assert_allclose(o["u"][-1], ou, rtol=0, atol=1e-13)
    np.testing.assert_allclose(o["v"][-1], ov, rtol=0, atol=1e-13)
    np.testing.assert_allclose(o["T"][-1], oT, rtol=0, atol=1e-13)
    for f, w in ((o["u"][0], wu), (o["v"][0], wv), (o["up"][0], wu), (o["vp"][0], wv)):
        np.testing.assert_allclose(f, w, rtol=0, atol=1e-15)


# ---- post-processing: stream function, vorticity, force coefficients ---------------------------

def test_stream_function_and_vorticity(golden):
    """solver.cpp:1784-1887 on the reference's step-0 and step-2 velocities, bit for bit
    (vorticity including the one-sided wall and outer rows)."""
    d, g = golden, _g(golden)
    n0, n1 = d["n0"], d["n1"]
    for st in ("s0", "s2"):
        si, vo = np.zeros((n1, n0)), np.zeros((n1, n0))
        lib().or_si_ref(n0, n1, d["deta"], d[st + "_u"], d[st + "_v"], g["xix"], g["xiy"], g["ajac"], si)
        lib().or_vort_ref(n0, n1, d["dxi"], d["deta"], d[st + "_u"], d[st + "_v"], g["xix"], g["xiy"], g["etx"],
                          g["ety"], vo)
        assert np.array_equal(si, d[st + "_si"]), st
        assert np.array_equal(vo, d[st + "_vort"]), st


def test_force_coefficients(golden):
    """CL, CD, CM, their pressure and vorticity parts and Nu from step 0's wall row, against
    the reference's COEFF_HIS.dat and COEFF_HIS_pr_vor.dat lines (printed with 8 decimals)."""
    d, g = golden, _g(golden)
    out = np.zeros(8)
    lib().or_forces_ref(d["n0"], d["n1"], d["dxi"], d["deta"], g["x"], g["y"], d["s0_p"], d["s0_vort"], d["s0_T"],
                        g["ajac"], g["etx"], g["ety"], d["Re"], 82.0, d["ar"], out)
    his, prv = d["s0_coeff_his_8dp"], d["s0_coeff_pr_vor_8dp"]
    np.testing.assert_allclose(out[[0, 1, 2, 7]], his[1:], rtol=0, atol=5e-9)
    np.testing.assert_allclose(out[3:7], prv[1:], rtol=0, atol=5e-9)
    # the wall columns of SURF_DIST.dat (6 significant digits): p and vorticity
    surf = d["s0_surf_dist_6sig"]
    np.testing.assert_allclose(surf[:, 1], d["s0_p"][0], rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(surf[:, 2], d["s0_vort"][0], rtol=1e-5, atol=1e-12)
