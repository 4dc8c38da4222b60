"""The benchmark configurations at their own sizes and physics on the GPU.

Config 2 (513^2 and 1025^2 to a 1e-8 residual, both precisions), and subsets of the
config-4 draw (uneven cases in one pack) and of the config-5 request stream (served
from one pack), checked through size-independent properties: convergence, agreement
of the mixed path with FP64, the steady heat balance, and every case computing the
same bits as itself alone.
"""
import numpy as np
import pytest

from cfd_solver_amd.grid import cavity_grid
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [513, 1025])
def test_config2_solves_to_1e8_and_mixed_matches_fp64(n):
    """Config 2 (skew 0.1, tanh 1.5, Re 400, Ri 1, Pr 0.71) from rest to mass and energy
    residual ratios of 1e-8 in both precisions within 70 FAS cycles (measured 55-57);
    the mixed fields within 1e-6 relative L2 of the FP64 ones (measured below 1e-12),
    Nu within 1e-6; the converged heat balance: with adiabatic sides the heat entering at
    the bottom leaves at the lid, so the two wall Nusselt numbers (same wall length)
    agree to the discretisation error (below 1e-3 relative)."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase
    g = cavity_grid(n, skew=0.1, beta=1.5)
    out = {}
    for prec in (CFD_FP64, CFD_MIXED):
        print(f"\n[gpu] config 2 at {n}x{n}, precision {prec}, to 1e-8", flush=True)
        c = CavityCase(g, precision=prec, Re=400.0, Ri=1.0, Pr=0.71)
        st, ok = c.solve(tol=1e-8, max_outer=70)
        print(f"[gpu] {st.iter} cycles, mass {st.mass_res:.2e}, energy {st.energy_res:.2e}, Nu {st.nu_hot:.8f} "
              f"/ {st.nu_cold:.8f}", flush=True)
        assert ok, (st.iter, st.mass_res, st.energy_res)
        out[prec] = (st, c.fields())
        c.close()
    (s64, f64), (s32, f32) = out[CFD_FP64], out[CFD_MIXED]
    for k in ("u", "v", "T"):
        assert rel_l2(f32[k], f64[k]) < 1e-6, k
    pa, pb = f32["p"] - f32["p"][1:-1, 1:-1].mean(), f64["p"] - f64["p"][1:-1, 1:-1].mean()
    assert rel_l2(pa, pb) < 1e-6
    assert abs(s32.nu_hot - s64.nu_hot) < 1e-6 * abs(s64.nu_hot)
    assert abs(s64.nu_hot - s64.nu_cold) < 1e-3 * abs(s64.nu_hot)


def test_config4_draw_in_one_pack_equals_cases_alone():
    """The first 40 cases of the config-4 draw of at most 257 nodes per side (Re 50-1000,
    Ri 0.1-10, Pr 0.71/7, skew 0-0.15) in one mixed-precision pack, 12 cycles: four of
    them, one of each size and both coarsest levels, are bitwise equal to themselves in
    a pack of their own, and every case's statistics are finite or it is flagged."""
    import ctypes as C
    from cfd_solver_amd._lib import CFD_MIXED, load
    from cfd_solver_amd.cavity import CavityPack, default_params
    from cfd_solver_amd.workloads import batched_cases
    L = load()
    cases = [c for c in batched_cases() if c[0] <= 257][:40]
    prm = default_params(precision=CFD_MIXED, graph=1)
    depth = [L.cfd_fas_coarsest(n, n, C.byref(default_params(**ph))) for n, _, ph in cases]
    sizes = [257, 129, 65, 33, 17]
    slots = [sum(1 for (n, _, _), d in zip(cases, depth) if d <= m <= n) for m in sizes]
    print(f"\n[gpu] 40 config-4 cases in a pack {dict(zip(sizes, slots))}, 12 cycles", flush=True)
    g0 = cavity_grid(257)
    pk = CavityPack(257, 257, g0.dxi, g0.deta, slots, prm)
    ids = [pk.add_case(cavity_grid(n, skew=sk), **ph) for n, sk, ph in cases]
    st = pk.step(12)
    assert all(st[i].iter == 12 for i in ids)
    picks = []
    for n in (257, 129, 65):
        for d in (17, 33):
            k = next((j for j, ((m, _, _), e) in enumerate(zip(cases, depth)) if m == n and e == d), None)
            if k is not None and len(picks) < 4:
                picks.append(k)
    assert len(picks) >= 3
    for k in picks:
        n, sk, ph = cases[k]
        g = cavity_grid(n, skew=sk)
        one = CavityPack(n, n, g.dxi, g.deta, [1] * (5 - sizes.index(n)), prm)
        cid = one.add_case(g, **ph)
        s1 = one.step(12)
        fa, fb = one.fields(cid), pk.fields(ids[k])
        for f in ("u", "v", "T", "p"):
            assert np.array_equal(fa[f], fb[f]), (k, n, f)
        assert s1[cid].mass_l1 == st[ids[k]].mass_l1
        one.close()
    pk.close()


def test_config5_stream_served_from_one_pack():
    """Ten config-5 requests (65-257 nodes per side, Re 50-1000, Ri 0.1-10, Pr 0.71/7)
    arriving over the first steps of a packed server with fewer slots than requests:
    every request finishes (converged, at its cap, or flagged non-finite), and two of
    them return exactly the fields and statistics of the same request served alone."""
    from cfd_solver_amd._lib import CFD_MIXED
    from cfd_solver_amd.cavity import default_params
    from cfd_solver_amd.serve import CavityServer
    from cfd_solver_amd.workloads import served_cases
    reqs = served_cases(10, 20.0, kmin=6, kmax=8, re=(50.0, 1000.0), ri=(0.1, 10.0), pr=(0.71, 7.0))
    g0 = cavity_grid(257)
    shape = (257, 257, g0.dxi, g0.deta, [2, 4, 6, 6, 6])
    prm = default_params(precision=CFD_MIXED, graph=1)
    print("\n[gpu] 10 config-5 requests on a packed server", flush=True)

    def serve(idx, staggered):
        srv = CavityServer(prm, pack=shape)
        tickets, res = {}, {}
        for step in range(100000):
            if staggered and step < 2 * len(idx):
                if step % 2 == 0:
                    _, n, sk, ph = reqs[idx[step // 2]]
                    tickets[srv.submit(cavity_grid(n, skew=sk), tol=1e-8, max_outer=40, **ph)] = idx[step // 2]
            elif not staggered and step == 0:
                for i in idx:
                    _, n, sk, ph = reqs[i]
                    tickets[srv.submit(cavity_grid(n, skew=sk), tol=1e-8, max_outer=40, **ph)] = i
            srv.step()
            for t in list(tickets):
                r = srv.poll(t, fields=True)
                if r is not None:
                    res[tickets.pop(t)] = r
            if not tickets and len(res) == len(idx):
                break
        srv.close()
        return res

    mix = serve(list(range(10)), True)
    assert all(mix[i]["state"] in ("done", "notconv", "failed") for i in range(10))
    print("  states:", [mix[i]["state"] for i in range(10)], flush=True)
    for i in (3, 7):
        alone = serve([i], False)[i]
        assert alone["state"] == mix[i]["state"] and alone["stats"].iter == mix[i]["stats"].iter
        for f in ("u", "v", "T", "p"):
            assert np.array_equal(alone["fields"][f], mix[i]["fields"][f]), (i, f)
