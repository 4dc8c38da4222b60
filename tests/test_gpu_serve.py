"""Serving: continuous batching gives every request the result it gets alone."""
import numpy as np
import pytest

from cfd_solver_amd.grid import cavity_grid

pytestmark = pytest.mark.gpu


def _drain(srv, tickets, cap=200000):
    res = {}
    for _ in range(cap):
        srv.step()
        for t in tickets:
            if t not in res:
                r = srv.poll(t, fields=True)
                if r is not None:
                    res[t] = r
        if len(res) == len(tickets):
            return res
    raise AssertionError("requests did not finish")


def test_served_alone_equals_in_mix():
    print("\n[gpu] a request served alone vs in a mix of shapes and physics", flush=True)
    from cfd_solver_amd.serve import CavityServer
    ga = cavity_grid(33, skew=0.05)
    pa = dict(Re=100.0, Ri=1.0, Pr=0.71)
    alone = CavityServer(slots_per_batch=4)
    ta = alone.submit(ga, tol=1e-5, max_outer=4000, **pa)
    ra = _drain(alone, [ta])[ta]
    alone.close()
    assert ra["state"] == "done" and ra["stats"].mass_res <= 1e-5

    mix = CavityServer(slots_per_batch=2)
    t1 = mix.submit(cavity_grid(17, skew=0.1), tol=1e-5, max_outer=4000, Re=100.0, Ri=0.5, Pr=7.0)
    t2 = mix.submit(cavity_grid(33, skew=0.02, beta=1.3), tol=1e-5, max_outer=4000, Re=800.0, Ri=0.2, Pr=0.71)
    for _ in range(7):          # the request under test arrives while others run
        mix.step()
    t3 = mix.submit(ga, tol=1e-5, max_outer=4000, **pa)
    t4 = mix.submit(cavity_grid(33, skew=0.12), tol=1e-5, max_outer=4000, Re=50.0, Ri=9.0, Pr=7.0)
    res = _drain(mix, [t1, t2, t3, t4])
    rb = res[t3]
    assert rb["state"] == "done"
    assert rb["stats"].iter == ra["stats"].iter and rb["stats"].mass_l1 == ra["stats"].mass_l1
    for k in ("u", "v", "T", "p"):
        assert np.array_equal(rb["fields"][k], ra["fields"][k]), k
    # the neighbours may converge, stop at their cap or diverge (a non-finite
    # residual retires the slot); none of it reaches the request under test
    print("  neighbour states:", [res[t]["state"] for t in (t1, t2, t4)], flush=True)
    assert all(res[t]["state"] in ("done", "notconv", "failed") for t in (t1, t2, t4))
    mix.close()


def test_server_budget_releases_empty_batches():
    print("\n[gpu] HBM budget admits one batch at a time", flush=True)
    from cfd_solver_amd.cavity import default_params
    from cfd_solver_amd.serve import CavityServer
    import ctypes as C
    from cfd_solver_amd import _lib
    prm = default_params()
    g17, g33 = cavity_grid(17), cavity_grid(33)
    n17, n33 = C.c_size_t(), C.c_size_t()
    _lib.call("cfd_batch_bytes_needed", 17, 17, g17.dxi, g17.deta, 2, C.byref(prm), C.byref(n17))
    _lib.call("cfd_batch_bytes_needed", 33, 33, g33.dxi, g33.deta, 2, C.byref(prm), C.byref(n33))
    budget = n33.value + n17.value // 2       # either batch fits, never both
    srv = CavityServer(prm, hbm_budget=budget, slots_per_batch=2)
    ts = [srv.submit(g17, tol=1e-4, max_outer=1500, Re=100.0), srv.submit(g33, tol=1e-4, max_outer=1500, Re=100.0),
          srv.submit(g17, tol=1e-4, max_outer=1500, Re=150.0)]
    seen = set()
    res = {}
    for _ in range(100000):
        srv.step()
        info = srv.info()
        assert info["batches"] <= 1 and info["bytes"] <= budget
        seen.add(info["batches"])
        for t in ts:
            if t not in res:
                r = srv.poll(t)
                if r is not None:
                    res[t] = r
        if len(res) == 3:
            break
    assert all(r["state"] in ("done", "notconv") for r in res.values())
    assert res[ts[1]]["state"] == "done"
    srv.close()


def test_server_iteration_cap():
    print("\n[gpu] max_outer stops a request", flush=True)
    from cfd_solver_amd.serve import CavityServer
    srv = CavityServer(slots_per_batch=1)
    t = srv.submit(cavity_grid(33), tol=1e-14, max_outer=5)
    r = _drain(srv, [t])[t]
    assert r["state"] == "notconv" and r["stats"].iter == 5
    srv.close()


def test_packed_server_alone_equals_in_mix():
    print("\n[gpu] packed server: a request alone vs arriving into a full mix of sizes", flush=True)
    from cfd_solver_amd.serve import CavityServer
    g0 = cavity_grid(129)
    shape = (129, 129, g0.dxi, g0.deta, [2, 3, 4, 2])   # slabs 129, 65, 33, 17
    ga = cavity_grid(65, skew=0.05)
    pa = dict(Re=100.0, Ri=1.0, Pr=0.71)
    alone = CavityServer(pack=shape)
    ta = alone.submit(ga, tol=1e-6, max_outer=300, **pa)
    ra = _drain(alone, [ta])[ta]
    alone.close()
    assert ra["state"] == "done" and ra["stats"].mass_res <= 1e-6

    mix = CavityServer(pack=shape)
    ts = [mix.submit(cavity_grid(129, skew=0.1), tol=1e-6, max_outer=60, Re=150.0, Ri=1.0, Pr=0.71),
          mix.submit(cavity_grid(129, skew=0.02, beta=1.3), tol=1e-6, max_outer=40, Re=600.0, Ri=0.2, Pr=7.0),
          mix.submit(cavity_grid(65, skew=0.12), tol=1e-6, max_outer=30, Re=80.0, Ri=4.0, Pr=7.0)]
    for _ in range(3):
        mix.step()
    assert mix.info()["batches"] == 1
    t3 = mix.submit(ga, tol=1e-6, max_outer=300, **pa)   # waits: the 65 slab is full
    t4 = mix.submit(cavity_grid(129), tol=1e-6, max_outer=20, Re=100.0)
    res = _drain(mix, ts + [t3, t4])
    rb = res[t3]
    assert rb["state"] == "done"
    assert rb["stats"].iter == ra["stats"].iter and rb["stats"].mass_l1 == ra["stats"].mass_l1
    for k in ("u", "v", "T", "p"):
        assert np.array_equal(rb["fields"][k], ra["fields"][k]), k
    print("  others:", [res[t]["state"] for t in ts + [t4]], flush=True)
    assert all(res[t]["state"] in ("done", "notconv", "failed") for t in ts + [t4])
    # a request of a size the pack has no slab for fails at once
    tb = mix.submit(cavity_grid(97), tol=1e-6, max_outer=10)
    mix.step()
    assert mix.poll(tb)["state"] == "failed"
    mix.close()
