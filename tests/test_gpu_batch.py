"""Batches of equally sized cases: every slot evolves exactly as the same case alone.

A slot shares each launch with the other slots but owns its arrays, scalars and
physics, so its fields must be bitwise equal to a single-case run of the same
case (which the oracle tests pin), whatever the other slots hold, and a slot
refilled in the middle of a run must not disturb the others.
"""
import numpy as np
import pytest

from cfd_solver_amd.grid import cavity_grid

pytestmark = pytest.mark.gpu

# one FAS depth for the mixed physics of a batch and for each case alone (the automatic
# coarsest level depends on Re, Ri and Pr)
DEPTH = dict(fas_min_n=33)

CASES = [  # (skew, beta, physics)
    (0.05, None, dict(Re=100.0, Ri=1.0, Pr=0.71)),
    (0.12, 1.5, dict(Re=400.0, Ri=0.3, Pr=7.0)),
    (0.0, 1.2, dict(Re=900.0, Ri=5.0, Pr=0.71, t_hot=2.0, lid_u=0.5)),
]


def _alone(n, skew, beta, phys, precision, steps, n1=None):
    from cfd_solver_amd.cavity import CavityCase, default_params
    c = CavityCase(cavity_grid(n, n1, skew=skew, beta=beta), default_params(precision=precision, **DEPTH, **phys))
    st = None
    for _ in range(steps):
        st = c.step()
    f = c.fields()
    c.close()
    return f, st


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_slots_equal_single_cases(precision):
    print(f"\n[gpu] batch of 4 slots (3 cases) vs single cases, precision {precision}", flush=True)
    from cfd_solver_amd.cavity import CavityBatch, default_params
    n, steps = 65, 40
    g0 = cavity_grid(n)
    b = CavityBatch(n, n, g0.dxi, g0.deta, 4, default_params(precision=precision, **DEPTH))
    for z, (skew, beta, phys) in enumerate(CASES):
        b.set_case(z if z < 1 else z + 1, cavity_grid(n, skew=skew, beta=beta), **phys)  # slot 1 stays empty
    for _ in range(steps):
        st = b.step()
    assert st[1].iter == 0
    for z, (skew, beta, phys) in enumerate(CASES):
        slot = z if z < 1 else z + 1
        fa, sa = _alone(n, skew, beta, phys, precision, steps)
        fb = b.fields(slot)
        for k in ("u", "v", "T", "p"):
            assert np.array_equal(fa[k], fb[k]), (slot, k, np.abs(fa[k] - fb[k]).max())
        assert st[slot].iter == steps and st[slot].p_iters == sa.p_iters
        assert st[slot].mass_l1 == sa.mass_l1 and st[slot].nu_hot == sa.nu_hot
    b.close()


def test_batch_refill_leaves_other_slots_alone():
    print("\n[gpu] refill one slot mid-run", flush=True)
    from cfd_solver_amd.cavity import CavityBatch, default_params
    n = 33
    g0 = cavity_grid(n)
    b = CavityBatch(n, n, g0.dxi, g0.deta, 2, default_params(**DEPTH))
    (s0, b0, p0), (s1, b1, p1), (s2, b2, p2) = CASES
    b.set_case(0, cavity_grid(n, skew=s0, beta=b0), **p0)
    b.set_case(1, cavity_grid(n, skew=s1, beta=b1), **p1)
    b.step(25)
    b.release(1)
    b.step(5)                      # slot 1 idle: slot 0 goes on alone
    b.set_case(1, cavity_grid(n, skew=s2, beta=b2), **p2)
    st = b.step(20)
    fa, _ = _alone(n, s0, b0, p0, 0, 50)
    fc, sc = _alone(n, s2, b2, p2, 0, 20)
    f0, f1 = b.fields(0), b.fields(1)
    for k in ("u", "v", "T", "p"):
        assert np.array_equal(f0[k], fa[k]), k
        assert np.array_equal(f1[k], fc[k]), k
    assert st[0].iter == 50 and st[1].iter == 20 and st[1].mass_l1 == sc.mass_l1
    b.close()


def test_batch_eager_equals_graph():
    print("\n[gpu] batch eager vs graph", flush=True)
    from cfd_solver_amd.cavity import CavityBatch, default_params
    n = 65
    g0 = cavity_grid(n)
    out = []
    for graph in (0, 1):
        b = CavityBatch(n, n, g0.dxi, g0.deta, 3, default_params(graph=graph, precision=1, **DEPTH))
        for z, (skew, beta, phys) in enumerate(CASES):
            b.set_case(z, cavity_grid(n, skew=skew, beta=beta), **phys)
        st = b.step(30)
        out.append(([b.fields(z) for z in range(3)], [(s.p_iters, s.mass_l1) for s in st]))
        b.close()
    (fe, se), (fg, sg) = out
    assert se == sg
    for z in range(3):
        for k in ("u", "v", "T", "p"):
            assert np.array_equal(fe[z][k], fg[z][k]), (z, k)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_slot_equals_alone_at_257(precision):
    """At 257x257 the finest multigrid level has 165 tiles per slot: the block size
    chosen for it (which fixes the order of its partial sums) must not depend on
    how many slots share the launch, or a slot would differ from the case alone."""
    print(f"\n[gpu] 257x257: 2-slot batch vs alone, precision {precision}", flush=True)
    from cfd_solver_amd.cavity import CavityBatch, default_params
    n, steps = 257, 4
    skew, beta, phys = CASES[1]
    g0 = cavity_grid(n)
    b = CavityBatch(n, n, g0.dxi, g0.deta, 2, default_params(precision=precision, **DEPTH))
    b.set_case(0, cavity_grid(n, skew=skew, beta=beta), **phys)
    b.set_case(1, cavity_grid(n, skew=CASES[0][0], beta=CASES[0][1]), **CASES[0][2])
    for _ in range(steps):
        st = b.step()
    fa, sa = _alone(n, skew, beta, phys, precision, steps)
    fb = b.fields(0)
    for k in ("u", "v", "T", "p"):
        assert np.array_equal(fa[k], fb[k]), (k, np.abs(fa[k] - fb[k]).max())
    assert st[0].p_iters == sa.p_iters and st[0].mass_l1 == sa.mass_l1
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_slots_equal_alone_on_rectangular_grid(precision):
    """97 x 41 nodes: the north wall row lies outside every tile (40 is a multiple
    of the tile height), so the edge tiles of each slot set it; slot 1 stays empty.
    Every filled slot is bitwise equal to its case alone."""
    print(f"\n[gpu] 97x41: 3-slot batch (slot 1 empty) vs alone, precision {precision}", flush=True)
    from cfd_solver_amd.cavity import CavityBatch, default_params
    n0, n1, steps = 97, 41, 4
    g0 = cavity_grid(n0, n1)
    b = CavityBatch(n0, n1, g0.dxi, g0.deta, 3, default_params(precision=precision, **DEPTH))
    for z, (skew, beta, phys) in ((0, CASES[1]), (2, CASES[2])):
        b.set_case(z, cavity_grid(n0, n1, skew=skew, beta=beta), **phys)
    for _ in range(steps):
        st = b.step()
    for z, (skew, beta, phys) in ((0, CASES[1]), (2, CASES[2])):
        fa, sa = _alone(n0, skew, beta, phys, precision, steps, n1=n1)
        fb = b.fields(z)
        for k in ("u", "v", "T", "p"):
            assert np.array_equal(fa[k], fb[k]), (z, k, np.abs(fa[k] - fb[k]).max())
        assert st[z].p_iters == sa.p_iters and st[z].mass_l1 == sa.mass_l1
    b.close()
