"""Inputs and expected outputs for the reference's outer-boundary stages
(solver.cpp:1351-1381, 1543-1655, 1731-1782), shared by the oracle and GPU tests.

Two sources:
  * the compiled reference's step-2 arrays (golden fixture): each stage is fed
    that step's inputs with its output rows cleared and must rebuild them;
  * a closed form on two polar rows (radii r0 < r1 on common rays): a free
    stream plus a radial term A/r^2 and a potential vortex G/r.  The far-field
    prediction vr0 cr^2 + vr_inf (1 - cr^2) reproduces A/r^2 exactly, and
    vth0 cr^k + vth_inf (1 - cr^k) reproduces G/r exactly for k = 1 (the branch a
    circulation above 1e-2 selects) and gives G r0 / r1^2 for k = 2; the wall row
    moves as a rigid rotation (-a y, a x); at CFL uet dt / deta = 1 the convected
    outflow temperature is the old profile shifted by one row.
"""
import math

import numpy as np


def uet_start(d):
    """Contravariant eta velocity at the start of step 2 (solver.cpp:1062-1068)."""
    return d["grid_etx"] * d["s2_uold"] + d["grid_ety"] * d["s2_vold"]


def farfield_golden(d):
    """Step-2 inputs of the far-field stage (its output rows cleared) and the expected arrays."""
    u, v, T = d["s2_u"].copy(), d["s2_v"].copy(), d["s2_T"].copy()
    up, vp = d["s2_up"].copy(), d["s2_vp"].copy()
    for a in (u, v):
        a[0] = 0.0
        a[-1] = 0.0
    T[-1] = 0.0
    up[0] = 0.0
    vp[0] = 0.0
    args = dict(x=d["grid_x"], y=d["grid_y"], u=u, v=v, T=T, up=up, vp=vp, Told=d["s2_Told"], uet=uet_start(d),
                etx=d["grid_etx"], ety=d["grid_ety"], ajac=d["grid_ajac"], xnox=d["grid_xnox"],
                xnoy=d["grid_xnoy"], uinf=d["uinf"], vinf=d["vinf"], dt=d["dt"], deta=d["deta"], speed_amp=0.0)
    want = dict(u=d["s2_u"], v=d["s2_v"], T=d["s2_T"], up=d["s2_up"], vp=d["s2_vp"])
    return args, want


def polar_case(G, A=0.3, amp=0.25, n0=33, n1=6, r0=5.0, r1=6.0):
    """Synthetic far-field inputs on rows n1-2 (radius r0) and n1-1 (radius r1).

    Returns (args, expected outer u, v, T, expected wall u, v, mask of the columns
    that take the prediction branch, decay exponent k)."""
    deta = 1.0 / (n1 - 1)
    dt = 0.01
    th = -2.0 * math.pi * np.arange(n0) / (n0 - 1)
    th[-1] = th[0]
    r = np.linspace(1.0, r0, n1 - 1).tolist() + [r1]
    R = np.array(r)[:, None] * np.ones((1, n0))
    x, y = R * np.cos(th)[None, :], R * np.sin(th)[None, :]
    uinf, vinf = math.sin(82 * math.pi / 180), math.cos(82 * math.pi / 180)
    c, s = np.cos(th)[None, :], np.sin(th)[None, :]
    vr = uinf * c + vinf * s + A / R**2
    vt = -uinf * s + vinf * c + G / R
    u = vr * c - vt * s
    v = vr * s + vt * c
    # metric terms that make the circulation integrand -v_theta (|J| = 1)
    etx, ety, ajac = c * np.ones_like(R), s * np.ones_like(R), np.ones_like(R)
    xnox, xnoy = -np.cos(th), -np.sin(th)  # inward normals: vnn < 0 where the stream leaves
    Told = 0.4 + 0.1 * np.arange(n1)[:, None] * np.ones((1, n0))
    uet = np.full((n1, n0), deta / dt)  # CFL 1
    circ_est = -np.mean(vt[n1 - 2, :-1]) * (n0 - 1) / (n0 - 2)
    k = 1 if abs(circ_est) > 1e-2 else 2
    vt1_vortex = G / r1 if k == 1 else G * r0 / r1**2
    cth, sth = np.cos(th), np.sin(th)
    vr1 = uinf * cth + vinf * sth + A / r1**2
    vt1 = -uinf * sth + vinf * cth + vt1_vortex
    out_u, out_v = vr1 * cth - vt1 * sth, vr1 * sth + vt1 * cth
    leave = uinf * xnox + vinf * xnoy < 0
    out_T = np.where(leave, Told[n1 - 2], 0.0)
    out_u, out_v = np.where(leave, out_u, uinf), np.where(leave, out_v, vinf)
    wall_u, wall_v = -amp * y[0], amp * x[0]
    zero = np.zeros((n1, n0))
    uu, vv = u.copy(), v.copy()
    uu[-1] = vv[-1] = 0.0
    args = dict(x=x, y=y, u=uu, v=vv, T=Told.copy(), up=zero.copy(), vp=zero.copy(), Told=Told, uet=uet, etx=etx,
                ety=ety, ajac=ajac, xnox=xnox, xnoy=xnoy, uinf=uinf, vinf=vinf, dt=dt, deta=deta, speed_amp=amp)
    return args, (out_u, out_v, out_T), (wall_u, wall_v), leave, k, circ_est


def exit_p_golden(d):
    p = d["s2_p"].copy()
    p[-1] = 0.0
    args = dict(u=d["s2_u"], v=d["s2_v"], T=d["s2_T"], uold=d["s2_uold"], vold=d["s2_vold"], p=p,
                xnox=d["grid_xnox"], xnoy=d["grid_xnoy"], uinf=d["uinf"], vinf=d["vinf"], dt=d["dt"], Re=d["Re"],
                Ri=float(d["s2_Ri"][0]))
    return args, d["s2_p"]


def outflow_golden(d):
    up, vp = d["s2_up"].copy(), d["s2_vp"].copy()
    up[-1] = 0.0
    vp[-1] = 0.0
    return dict(up=up, vp=vp, u=d["s2_uold"], v=d["s2_vold"], xnox=d["grid_xnox"], xnoy=d["grid_xnoy"],
                uinf=d["uinf"], vinf=d["vinf"]), (d["s2_up"], d["s2_vp"])


def golden_step_inputs(d):
    """Grid, operators, state at the end of step 1 and parameters of step 2."""
    g = {k: d["grid_" + k] for k in ["x", "y", "xix", "xiy", "etx", "ety", "alph", "beta", "gamma", "ajac",
                                     "xnox", "xnoy"]}
    ops = {k: d[k] for k in ["AU", "AT", "BU", "BT", "A9"]}
    st = dict(u=d["s2_uold"], v=d["s2_vold"], T=d["s2_Told"], p=d["s1_p"], up=d["s1_up"], vp=d["s1_vp"])
    prm = dict(dxi=d["dxi"], deta=d["deta"], dt=d["dt"], Re=d["Re"], Pr=d["Pr"], Ri=float(d["s2_Ri"][0]),
               uinf=d["uinf"], vinf=d["vinf"])
    return g, ops, st, prm
