"""The C ABI: the library loads without a GPU and exports every declared symbol."""
import os
import re

from cfd_solver_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "cfd_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cfd_\w+)\s*\(", txt, flags=re.M)))


def test_header_symbols_exported():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_version_and_device_count_without_gpu():
    import ctypes as C
    lib = _lib.load()
    assert lib.cfd_version().startswith(b"cfd_gpu")
    n = C.c_int(-1)
    assert lib.cfd_device_count(C.byref(n)) == 0
    assert n.value >= 0


def test_argument_errors_need_no_gpu():
    import numpy as np
    lib = _lib.load()
    z = np.zeros((3, 3))
    # grids below 5 x 5 are rejected before any device work
    assert lib.cfd_ref_assemble_p9(3, 3, 0.5, 0.5, z, z, z, z, np.zeros((9, 3, 3))) == -1
    # the reference-mode case checks its grid and parameters before touching the device
    import ctypes as C
    from cfd_solver_amd.refcase import ref_params
    h = C.c_void_p()
    n = np.zeros(3)
    assert lib.cfd_ref_case_create(3, 3, 0.5, 0.5, *[z] * 10, None, None, n, n, C.byref(ref_params()),
                                   C.byref(h)) == -1 and not h.value
    z7 = np.zeros((7, 7))
    n7 = np.zeros(7)
    assert lib.cfd_ref_case_create(7, 7, 0.5, 0.5, *[z7] * 10, None, None, n7, n7,
                                   C.byref(ref_params(p_maxit=0)), C.byref(h)) == -1
    assert lib.cfd_ref_case_step(None, 1, None) == -1


def test_batch_argument_errors_need_no_gpu():
    import ctypes as C
    lib = _lib.load()
    prm = _lib.Params()
    lib.cfd_params_default(C.byref(prm))
    h = C.c_void_p()
    # zero slots, tiny grids and a bad precision are rejected before any device work
    assert lib.cfd_batch_create(65, 65, 0.1, 0.1, 0, C.byref(prm), C.byref(h)) == -1
    assert lib.cfd_batch_create(4, 65, 0.1, 0.1, 2, C.byref(prm), C.byref(h)) == -1
    prm.precision = 7
    assert lib.cfd_batch_create(65, 65, 0.1, 0.1, 2, C.byref(prm), C.byref(h)) == -1
    # a case's own p' solve needs at least one BiCGStab iteration (the one-V-cycle
    # solve, p_maxit -1, is for its coarse levels: fas_coarse_pmaxit)
    lib.cfd_params_default(C.byref(prm))
    assert prm.fas_coarse_pmaxit == -1
    prm.p_maxit = -1
    assert lib.cfd_batch_create(65, 65, 0.1, 0.1, 2, C.byref(prm), C.byref(h)) == -1
    assert lib.cfd_batch_release(None, 0) == -1



def test_server_queue_needs_no_gpu():
    """Submitting and polling only queue host copies; no device work happens."""
    import ctypes as C
    import numpy as np
    from cfd_solver_amd.grid import cavity_grid
    from cfd_solver_amd.serve import CavityServer
    srv = CavityServer(hbm_budget=1 << 30, slots_per_batch=2)
    t = srv.submit(cavity_grid(17), tol=1e-6, max_outer=10, Re=50.0)
    assert srv.poll(t) is None
    assert srv.info() == dict(batches=0, bytes=0, finished=0)
    lib = _lib.load()
    st = C.c_int()
    assert lib.cfd_server_poll(srv._h, C.c_long(t + 99), C.byref(st), None, None, None, None, None, None) == -1
    srv.close()


def test_pack_shapes_need_no_gpu():
    """A pack's byte count comes from a dry build (no device work); chains that do not
    halve, fewer than two slabs and negative slot counts are rejected."""
    import ctypes as C
    from cfd_solver_amd.cavity import CavityPack, default_params
    lib = _lib.load()
    prm = default_params(precision=1)
    b = CavityPack.bytes_needed(129, 129, 1 / 128, 1 / 128, [2, 4, 4, 2], prm)
    b1 = CavityPack.bytes_needed(129, 129, 1 / 128, 1 / 128, [1, 4, 4, 2], prm)
    one = C.c_size_t()
    lib.cfd_batch_bytes_needed(129, 129, 1 / 128, 1 / 128, 1, C.byref(default_params(precision=1, fas_levels=1)),
                               C.byref(one))
    # one more 129^2 slot: a single-level batch slot plus the multigrid transfer arrays
    assert b1 > 0 and one.value <= b - b1 < 1.5 * one.value
    h = C.c_void_p()
    for n, slots in ((128, [1, 1]), (129, [1]), (129, [1, -1]), (9, [1, 1, 1])):
        arr = (C.c_int * len(slots))(*slots)
        assert lib.cfd_pack_create(n, n, 0.01, 0.01, len(slots), arr, C.byref(prm), C.byref(h)) == -1
    # a level of more than 2^28 nodes is refused (node offsets within a slot are 32-bit)
    big = C.c_size_t()
    arr = (C.c_int * 2)(1, 1)
    assert lib.cfd_pack_bytes_needed(16385, 16385, 1e-4, 1e-4, 2, arr, C.byref(prm), C.byref(big)) == -1
    assert lib.cfd_batch_bytes_needed(16385, 16385, 1e-4, 1e-4, 1, C.byref(prm), C.byref(big)) == -1
    assert lib.cfd_batch_bytes_needed(8193, 16385, 1e-4, 1e-4, 1, C.byref(prm), C.byref(big)) == 0 and big.value > 0
    assert lib.cfd_pack_release(None, 0) == -1 and lib.cfd_pack_launch(None) == -1
