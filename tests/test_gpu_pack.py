"""Packs: cases of uneven size share every launch and compute what they compute alone.

A pack's slab of one grid size holds the finest grids of the cases of that size
beside the coarse multigrid levels of every larger case; each launch of a stage
serves all of them.  Every case must still evolve bit for bit as the same case in a
batch of its own (automatic multigrid depth from its physics), while other cases
run, stop and are replaced around it.
"""
import numpy as np
import pytest

from cfd_solver_amd.grid import cavity_grid

pytestmark = pytest.mark.gpu

# (n, skew, beta, physics): coarsest levels 17 (A, D) and 33 (B, C)
CASES = {
    "A": (129, 0.05, None, dict(Re=100.0, Ri=1.0, Pr=0.71)),
    "B": (129, 0.12, 1.5, dict(Re=400.0, Ri=0.3, Pr=7.0)),
    "C": (65, 0.0, 1.2, dict(Re=900.0, Ri=5.0, Pr=0.71, t_hot=2.0, lid_u=0.5)),
    "D": (65, 0.1, None, dict(Re=200.0, Ri=0.5, Pr=0.71)),
}
SLOTS = [3, 5, 5, 3]  # 129, 65, 33, 17


def _alone(name, precision, steps):
    from cfd_solver_amd.cavity import CavityCase, default_params
    n, skew, beta, phys = CASES[name]
    c = CavityCase(cavity_grid(n, skew=skew, beta=beta), default_params(precision=precision, **phys))
    st = None
    for _ in range(steps):
        st = c.step()
    f = c.fields()
    c.close()
    return f, st


def _pack(precision, graph):
    from cfd_solver_amd.cavity import CavityPack, default_params
    g0 = cavity_grid(129)
    return CavityPack(129, 129, g0.dxi, g0.deta, SLOTS, default_params(precision=precision, graph=graph))


def _add(pk, name):
    n, skew, beta, phys = CASES[name]
    return pk.add_case(cavity_grid(n, skew=skew, beta=beta), **phys)


@pytest.mark.parametrize("precision,graph", [(0, 0), (1, 1)])
def test_pack_cases_equal_single_cases(precision, graph):
    print(f"\n[gpu] pack 129/65/33/17 of 4 cases vs single cases, precision {precision} graph {graph}", flush=True)
    from cfd_solver_amd.cavity import default_params
    from cfd_solver_amd import _lib
    import ctypes as C
    L = _lib.load()
    depth = {k: L.cfd_fas_coarsest(n, n, C.byref(default_params(**ph))) for k, (n, _, _, ph) in CASES.items()}
    assert depth == dict(A=17, B=33, C=33, D=17)
    steps = 12
    pk = _pack(precision, graph)
    ids = {k: _add(pk, k) for k in CASES}
    for _ in range(steps):
        st = pk.step()
    for k, cid in ids.items():
        fa, sa = _alone(k, precision, steps)
        fp = pk.fields(cid)
        for f in ("u", "v", "T", "p"):
            assert np.array_equal(fa[f], fp[f]), (k, f, np.abs(fa[f] - fp[f]).max())
        assert st[cid].iter == steps and st[cid].p_iters == sa.p_iters
        assert st[cid].mass_l1 == sa.mass_l1 and st[cid].nu_hot == sa.nu_hot and st[cid].mass_res == sa.mass_res
    for z in range(pk.capacity):
        if z not in ids.values():
            assert st[z].iter == 0
    pk.close()


def test_pack_release_and_refill():
    print("\n[gpu] pack: release a case mid-run, refill its slots with another size's chain", flush=True)
    pk = _pack(0, 1)
    a, b = _add(pk, "A"), _add(pk, "C")
    pk.step(6)
    pk.release(b)
    pk.step(2)                     # A goes on alone
    d = _add(pk, "D")              # takes C's freed 65/33 slots and a 17 slot
    st = pk.step(5)
    fa, sa = _alone("A", 0, 13)
    fd, sd = _alone("D", 0, 5)
    f0, f1 = pk.fields(a), pk.fields(d)
    for k in ("u", "v", "T", "p"):
        assert np.array_equal(f0[k], fa[k]), k
        assert np.array_equal(f1[k], fd[k]), k
    assert st[a].iter == 13 and st[d].iter == 5 and st[d].mass_l1 == sd.mass_l1
    pk.close()


def test_pack_full_and_unsupported_cases():
    print("\n[gpu] pack: a full slab and a case the pack cannot hold", flush=True)
    from cfd_solver_amd._lib import CfdError
    from cfd_solver_amd.cavity import CavityPack, default_params
    g0 = cavity_grid(129)
    pk = CavityPack(129, 129, g0.dxi, g0.deta, [1, 1, 1, 1], default_params())
    _add(pk, "A")
    with pytest.raises(CfdError) as e:
        _add(pk, "D")              # its 65 slot is taken by A's second level
    assert e.value.code == -7
    with pytest.raises(CfdError) as e:
        pk.add_case(cavity_grid(33), Re=100.0)   # a small grid cannot be a case's finest level here
    assert e.value.code == -1
    with pytest.raises(CfdError) as e:
        pk.add_case(cavity_grid(97), Re=100.0)   # no slab of that size
    assert e.value.code == -1
    pk.close()
