"""Server admission decisions that need no GPU work (run on the CPU)."""
import ctypes as C

from cfd_solver_amd.grid import cavity_grid


def test_request_larger_than_the_budget_fails_at_once():
    """A request whose batch alone needs more than the HBM budget fails in the first
    scheduling round instead of waiting forever (no batch is built, nothing launched)."""
    from cfd_solver_amd import _lib
    from cfd_solver_amd.cavity import default_params
    from cfd_solver_amd.serve import CavityServer
    prm = default_params()
    g = cavity_grid(33)
    need = C.c_size_t()
    _lib.call("cfd_batch_bytes_needed", 33, 33, g.dxi, g.deta, 2, C.byref(prm), C.byref(need))
    srv = CavityServer(prm, hbm_budget=need.value // 2, slots_per_batch=2)
    t = srv.submit(g, tol=1e-8, max_outer=10, Re=100.0)
    run, queued = srv.step()
    assert run == 0 and queued == 0
    r = srv.poll(t)
    assert r is not None and r["state"] == "failed"
    assert srv.info()["batches"] == 0
    srv.close()


def test_fas_coarsest_follows_the_reynolds_number():
    """Automatic multigrid depth: the coarsest level keeps Re / (n - 1) <= 32 and has at
    least 17 nodes per side, 33 when Pr > 1 or Ra = |Ri| Re^2 Pr > 2e5; an explicit
    fas_min_n overrides it; a grid that does not coarsen keeps its size."""
    from cfd_solver_amd import _lib
    from cfd_solver_amd.cavity import default_params
    L = _lib.load()
    co = lambda n, **kw: L.cfd_fas_coarsest(n, n, C.byref(default_params(**kw)))
    assert co(1025, Re=100.0) == 17 and co(1025, Re=400.0) == 17  # Ri 1, Pr 0.71: Ra <= 1.2e5
    assert co(1025, Re=400.0, Pr=7.0) == 33 and co(1025, Re=400.0, Ri=3.0) == 33  # Ra 3.4e5
    assert co(1025, Re=100.0, Ri=-1.0) == 17 and co(1025, Re=520.0, Ri=0.1) == 33  # Re / 16 > 32
    assert co(1025, Re=1000.0) == 33 and co(1025, Re=1100.0) == 65 and co(1025, Re=2100.0) == 129
    assert co(1025, Re=400.0, fas_min_n=65) == 65 and co(1025, Re=400.0, fas_min_n=17) == 17
    assert co(65, Re=1000.0) == 33 and co(33, Re=1000.0) == 33 and co(96, Re=100.0) == 96
    assert co(1025, Re=400.0, fas_levels=1) == 1025
