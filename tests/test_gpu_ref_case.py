"""The reference's whole O-grid time step on the GPU (cfd_ref_case_*) against the oracle.

The oracle composes the reference's stages in its order (pinned bit for bit to
the compiled reference's step 2 by test_oracle_golden.py); here its Gauss-Seidel
and SIP solves are repeated until converged, and the GPU solves the same systems
with BiCGStab to 1e-13, so the two agree to FP64 rounding of converged solves
(1e-9 relative L2 per field).
"""
import types

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "T", "p", "up", "vp")


def _grid(d):
    g = types.SimpleNamespace(n0=d["n0"], n1=d["n1"], dxi=d["dxi"], deta=d["deta"])
    for k in ["x", "y", "xix", "xiy", "etx", "ety", "alph", "beta", "gamma", "ajac", "xnox", "xnoy"]:
        setattr(g, k, d["grid_" + k])
    return g


def _prm_kw(d, **kw):
    return dict(Re=d["Re"], Pr=d["Pr"], dt=d["dt"], Ri=float(d["s2_Ri"][0]), uinf=d["uinf"], vinf=d["vinf"],
                mom_rtol=1e-13, p_rtol=1e-13, ar=d["ar"], **kw)


def _gpu_case(d, **kw):
    from cfd_solver_amd.refcase import RefCase, ref_params
    c = RefCase(_grid(d), ref_params(**_prm_kw(d, **kw)))
    c.set_state(d["s2_uold"], d["s2_vold"], d["s2_Told"], d["s1_p"], d["s1_up"], d["s1_vp"])
    return c


@pytest.mark.parametrize("nsteps", [1, 3])
def test_steps_match_oracle_with_converged_solves(golden, nsteps):
    from oracle.oracle import ref_step
    from tests.ref_bc_cases import golden_step_inputs
    print(f"\n[gpu] reference O-grid {golden['n0']}x{golden['n1']}, {nsteps} device time steps", flush=True)
    c = _gpu_case(golden)
    st = c.step(nsteps)
    g, ops, s, prm = golden_step_inputs(golden)
    for _ in range(nsteps):
        o = ref_step(g, ops, s, prm, inner_calls=12)
        s = {k: o[k] for k in FIELDS}
    f = c.state()
    for k in FIELDS:
        assert rel_l2(f[k], o[k]) < 1e-9, (k, rel_l2(f[k], o[k]))
    assert abs(st["dmax"] - o["dmax"]) <= 1e-8 * abs(o["dmax"])
    assert abs(st["circ"] - o["circ"]) <= 1e-8 * abs(o["circ"]) + 1e-15
    assert max(st["relres"]) <= 1e-13 and min(st["iters"]) >= 1
    # the step's post-processing (solver.cpp:1784-1982) against the oracle's on its own step
    import ctypes as C
    from oracle.oracle import lib
    d, n0, n1 = golden, golden["n0"], golden["n1"]
    gg = {k: np.ascontiguousarray(d["grid_" + k]) for k in ("x", "y", "xix", "xiy", "etx", "ety", "ajac")}
    si, vo, coef = np.zeros((n1, n0)), np.zeros((n1, n0)), np.zeros(8)
    lib().or_si_ref(n0, n1, d["deta"], o["u"], o["v"], gg["xix"], gg["xiy"], gg["ajac"], si)
    lib().or_vort_ref(n0, n1, d["dxi"], d["deta"], o["u"], o["v"], gg["xix"], gg["xiy"], gg["etx"], gg["ety"], vo)
    lib().or_forces_ref(n0, n1, d["dxi"], d["deta"], gg["x"], gg["y"], o["p"], vo, o["T"], gg["ajac"], gg["etx"],
                        gg["ety"], d["Re"], 82.0, d["ar"], coef)
    gsi, gvo = c.post()
    assert rel_l2(gsi, si) < 1e-9 and rel_l2(gvo, vo) < 1e-9
    for k, q in (("cl", 0), ("cd", 1), ("cm", 2), ("nu", 7)):
        assert abs(st[k] - coef[q]) <= 1e-8 * abs(coef[q]) + 1e-12, (k, st[k], coef[q])
    c.close()


def test_step_with_oscillating_cylinder(golden):
    """speed_amp != 0: the wall row moves as (-a y, a x) and the step stays equal to the oracle's."""
    from oracle.oracle import ref_step
    from tests.ref_bc_cases import golden_step_inputs
    print("\n[gpu] reference O-grid, cylinder speed amplitude 0.05", flush=True)
    c = _gpu_case(golden, speed_amp=0.05)
    c.step(1)
    g, ops, s, prm = golden_step_inputs(golden)
    o = ref_step(g, ops, s, dict(prm, speed_amp=0.05), inner_calls=12)
    f = c.state()
    for k in FIELDS:
        assert rel_l2(f[k], o[k]) < 1e-9, (k, rel_l2(f[k], o[k]))
    np.testing.assert_allclose(f["u"][0], -0.05 * golden["grid_y"][0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(f["v"][0], 0.05 * golden["grid_x"][0], rtol=0, atol=1e-15)
    c.close()


def test_time_accurate_oscillating_cylinder(golden):
    """time_accurate: every step advances time by dt first, so the cylinder's wall velocity
    (-a cos(2 pi F t) y, a cos(2 pi F t) x) and the wall pressure's local acceleration
    (accn_amp sin(2 pi F t)) change from step to step; three steps against the oracle's
    step taking the same times."""
    import math
    from oracle.oracle import ref_step
    from tests.ref_bc_cases import golden_step_inputs
    print("\n[gpu] reference O-grid, time-accurate, F 200, 3 steps", flush=True)
    F, amp = 200.0, 0.05
    acc = 2.0 * math.pi * F * amp
    c = _gpu_case(golden, speed_amp=amp, accn_amp=acc, F=F, time_accurate=1)
    c.step(3)
    g, ops, s, prm = golden_step_inputs(golden)
    for k in range(1, 4):
        o = ref_step(g, ops, s, dict(prm, speed_amp=amp, accn_amp=acc, F=F, time=k * prm["dt"]), inner_calls=12)
        s = {q: o[q] for q in FIELDS}
    f = c.state()
    for q in FIELDS:
        assert rel_l2(f[q], o[q]) < 1e-9, (q, rel_l2(f[q], o[q]))
    t = c.time()
    assert abs(t - 3 * golden["dt"]) < 1e-18
    cw = math.cos(2.0 * math.acos(-1.0) * F * t)
    np.testing.assert_allclose(f["u"][0], -amp * cw * golden["grid_y"][0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(f["v"][0], amp * cw * golden["grid_x"][0], rtol=0, atol=1e-15)
    c.close()


def test_restart_continues_bit_for_bit(golden, tmp_path):
    """Two steps, a restart file, a new case from it and one more step equal three steps
    of one case bit for bit; the output files hold the step's coefficients."""
    from cfd_solver_amd.refcase import RefCase
    print("\n[gpu] reference O-grid, restart after 2 of 3 steps", flush=True)
    a = _gpu_case(golden)
    a.step(3)
    b = _gpu_case(golden)
    b.step(2)
    path = tmp_path / "spa100.dat"
    b.save_restart(path, 2)
    c = RefCase.from_restart(_grid(golden), path, **_prm_kw(golden))
    st = c.step(1)
    fa, fc = a.state(), c.state()
    for q in FIELDS:
        assert np.array_equal(fa[q], fc[q]), q
    c.write_outputs(tmp_path, 3)
    his = (tmp_path / "COEFF_HIS.dat").read_text().split()
    assert his[0] == "0" and abs(float(his[1]) - st["cl"]) <= 5e-9 and abs(float(his[4]) - st["nu"]) <= 5e-9
    assert (tmp_path / "spt100.dat").read_text().startswith(f"zone\nI={golden['n0']}\nJ={golden['n1']}\n")
    assert len((tmp_path / "SURF_DIST.dat").read_text().splitlines()) == golden["n0"]
    for x in (a, b, c):
        x.close()
