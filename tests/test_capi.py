"""The C-ABI library loads and exports every entry point include/sadmrg.h declares
(no device calls: these run on the CPU-only CI host)."""
import ctypes
import os
import subprocess

import sadmrg_testlib as tl


def test_library_exports_every_declared_symbol():
    sad = tl.load_pkg()
    assert os.path.exists(sad.LIB_PATH), "libsadmrg.so must be built"
    lib = ctypes.CDLL(sad.LIB_PATH)
    declared = sad.exported_symbols()
    assert len(declared) >= 50
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    sad = tl.load_pkg()
    L = sad.lib()
    for n in sad.exported_symbols():
        assert getattr(L, n).restype is not None or n.endswith("_error")


def test_no_torch_or_cpp_types_in_header():
    import re
    text = open(os.path.join(tl.ROOT, "include", "sadmrg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for bad in ("torch", "std::", "template", "class "):
        assert bad not in text


def test_host_only_entry_points():
    """CGC coefficients and error reporting need no device."""
    sad = tl.load_pkg()
    assert abs(sad.clebsch_gordan(1, 1, 1, -1, 0, 0) - 2 ** -0.5) < 1e-15
    assert abs(sad.clebsch_gordan(2, 2, 2, -2, 0, 0) - 3 ** -0.5) < 1e-15
    c = ctypes.c_void_p()
    rc = sad.lib().sadmrg_tensor_destroy(None)
    assert rc == 0
    rc = sad.lib().sadmrg_ctx_sync(None)
    assert rc == sad.SadmrgError(1, "").code
    assert b"null argument" in sad.lib().sadmrg_last_error()


def test_cgc_tables_match_reference_fixture():
    """The engine's recoupling tables reproduce every tabulated reference entry."""
    import numpy as np
    sad = tl.load_pkg()
    z = np.load(os.path.join(tl.GOLDEN, "cgc_su2.npz"))
    M = int(z["max_angm"])
    cg = sad.Cgc(1)
    hv = z["hamilt_vec"]
    for i in range(M):
        for j in range(0, 7):
            for k in range(M):
                assert abs(cg.hamilt_vec(i, j, k) - hv[i, j, k]) <= 1e-12 * max(1, abs(hv[i, j, k]))
    for side, name in enumerate(["onsite_left", "onsite_rght"]):
        for idx, val in zip(z[name + "_idx"], z[name + "_val"]):
            j, m, l, i, p = np.unravel_index(idx, (M, M, M, M, 9))
            assert abs(cg.onsite(side, j, m, l, i, p) - val) <= 1e-12 * max(1, abs(val)), (name, j, m, l, i, p)
    for side, name in enumerate(["pairup_left", "pairup_rght"]):
        for idx, val in zip(z[name + "_idx"], z[name + "_val"]):
            j, i, k, m, p = np.unravel_index(idx, (M, M, M, M, 9))
            assert abs(cg.pairup(side, j, i, k, m, p) - val) <= 1e-12 * max(1, abs(val)), (name, j, i, k, m, p)
    for side, name in enumerate(["transformation_left", "transformation_rght"]):
        for idx, val in zip(z[name + "_idx"], z[name + "_val"]):
            j, m, l, i, k, p = np.unravel_index(idx, (M, M, M, M, M, 3))
            assert abs(cg.transformation(side, j, m, l, i, k, p) - val) <= 1e-12 * max(1, abs(val))
    for row in z["cg"]:
        j1, m1, j2, m2, j3, m3 = (int(x) for x in row[:6])
        assert abs(sad.clebsch_gordan(j1, m1, j2, m2, j3, m3) - row[6]) <= 1e-14


def test_cgc_zero_entries_stay_zero():
    """Entries absent from the reference table (zero) are zero in the engine."""
    import numpy as np
    sad = tl.load_pkg()
    z = np.load(os.path.join(tl.GOLDEN, "cgc_su2.npz"))
    M = int(z["max_angm"])
    cg = sad.Cgc(1)
    nz = set(z["transformation_left_idx"].tolist())
    rng = np.random.default_rng(3)
    for _ in range(3000):
        j, m, l, i, k = rng.integers(0, 12, size=5)
        p = int(rng.integers(0, 3))
        idx = np.ravel_multi_index((j, m, l, i, k, p), (M, M, M, M, M, 3))
        if idx not in nz:
            assert cg.transformation(0, j, m, l, i, k, p) == 0.0
