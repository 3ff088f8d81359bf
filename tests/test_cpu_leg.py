"""The CPU-baseline child process of bench.py (oracle/ref_centre_step.py) on a
small state the reference library itself converges (SU(2) J1-J2 3x4
cylinder, D = 32): tensors saved in the bench's naming, then timed in both
BLAS settings.  Its H_eff output equals the reference's in-process
application, its truncation keeps the whole weight of the normalised
state, and the environment updates run."""
import os
import subprocess
import sys

import numpy as np
import pytest

import sadmrg_testlib as tl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_centre_step_child(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)  # the reference's sweep writes its checkpoint files into the working directory
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refbind as rb
    if not rb.available():
        pytest.skip("oracle/_ref not built")
    sad = tl.load_pkg()
    width, length, D = 3, 4, 32
    ns = width * length
    rb.init(physpn=1, max_angm=20, max_dcut=D, psize=2)
    d = rb.RefDMRG(1, 0, length, width, D, 1, 0)
    hmap = sad.j1j2_cylinder(width, length)
    coup = np.zeros(ns + 1)
    coup[1], coup[2] = 1.0, 0.5
    d.set_couplings(hmap, coup)
    d.idmrg()
    d.sweep()
    il, ir = ns // 2 - 1, ns // 2
    tens = {"psi": tl.from_ref(d.input_vector(il, ir))}

    def put(slot, i, j=0):
        t = tl.from_ref(d.tensor(slot, i, j))
        if t["present"].any():
            tens[f"{slot}_{i}_{j}"] = t

    for i in (il, ir):
        put("uu", i)
    for i in (il - 1, il, ir, ir + 1):
        put("hh", i)
        for j in range(ns):
            put("opr", i, j)
    path = str(tmp_path / "centre.npz")
    tl.save_tensors(path, tens, params=np.asarray([width, length, D, il, ir]), hmap=np.asarray(hmap, np.int32),
                    coup=coup)
    want = tl.from_ref(d.heff(il, ir, tl.to_ref(tens["psi"], 1)))
    for blas in ("omp", "mkl"):
        out = str(tmp_path / f"heff_{blas}.npz")
        r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "ref_centre_step.py"), "--workload", "real",
                            "--blas", blas, "--threads", "2", "--budget", "0.5", "--tensors", path, "--out", out],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        res = __import__("json").loads(r.stdout.strip().splitlines()[-1])
        assert res["blas"] == blas and res["heff_s"] > 0 and res["heff_reps"] >= 1
        assert res["env_left_s"] > 0 and res["env_right_s"] > 0 and res["svd_s"] > 0
        assert abs(res["svd_kept_sum_sq"] - 1.0) < 1e-10, res
        got, _ = tl.load_tensors(out)
        assert tl.compare(got["heff"], want)[0] <= 1e-14, blas
