#!/usr/bin/env python3
"""CPU baseline runner (TEST INFRASTRUCTURE: only bench.py's cpu_baseline leg
runs it, as a child process).  Times the reference library (oracle/_ref, the
reference's own sources compiled here) on the host cores, in one BLAS
threading setting per process -- MKL fixes its threading layer at the first
BLAS call of a process:

  omp  sequential MKL, the reference's OpenMP over bond terms with `threads`
       threads (psize; dmrg_inst.cpp:283-371, block_tensor_inst.cpp:501-637);
  mkl  MKL's GNU-OpenMP layer with `threads` BLAS/LAPACK threads, psize 1.

Workloads:
  synthetic  the rotor-histogram centre bond that bench.py measures
             (sadmrg_testlib.build_workload): H_eff applications only;
  real       tensors of a warmed-up 6 x W cylinder saved by bench.py (an
             .npz of packed tensors): H_eff applications
             (hamiltonian_vector_multiplication_idmrg), the two truncated-
             basis environment updates of the step
             (prepare_site_operator_from_left(il), ..._from_right(ir);
             dmrg_inst.cpp:588-700) and the truncation
             (BlockTensor::svd with (p1, p2) = (1, 0), block_tensor_inst.cpp:501).

Prints one JSON line; with --out, also writes the reference H_eff output."""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "tests"))


def _timed(fn):
    t0 = time.perf_counter()
    r = fn()
    return time.perf_counter() - t0, r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["synthetic", "real"], required=True)
    ap.add_argument("--blas", choices=["omp", "mkl"], required=True)
    ap.add_argument("--threads", type=int, required=True)
    ap.add_argument("--budget", type=float, default=8.0, help="seconds of H_eff applications")
    ap.add_argument("--tensors", help="real: the .npz written by bench.py")
    ap.add_argument("--D", type=int, default=2048)
    ap.add_argument("--no-step", action="store_true", help="real: H_eff only (no truncation, no updates)")
    ap.add_argument("--no-svd", action="store_true",
                    help="real: no truncation (its sectors run in sequence, so only MKL's threads speed it up)")
    ap.add_argument("--out", help="write the reference H_eff output here (.npz)")
    a = ap.parse_args()

    import refbind as rb
    import sadmrg_testlib as tl
    if not rb.available():
        print(json.dumps({"error": "oracle/_ref not built"}))
        return
    psize = a.threads if a.blas == "omp" else 1
    rb.lib().ref_set_blas_threading(1 if a.blas == "omp" else 3, 1 if a.blas == "omp" else a.threads)
    res = {"workload": a.workload, "blas": a.blas, "threads": a.threads, "psize": psize}

    if a.workload == "synthetic":
        wl = tl.build_workload(D=a.D)
        d, il, ir = tl.reference_heff_setup(rb, wl, psize)
        vin = tl.to_ref(wl["psi"], 1)
    else:
        tens, extra = tl.load_tensors(a.tensors)
        width, length, D, il, ir = (int(x) for x in extra["params"])
        rb.init(physpn=1, max_angm=20, max_dcut=D, psize=psize)
        d = rb.RefDMRG(1, 0, length, width, D, 1, 0)
        d.set_couplings(np.asarray(extra["hmap"], np.int32), np.asarray(extra["coup"], np.float64))
        # names: "<slot>_<i>_<j>" for uu / hh / opr, and "psi"
        keep = []
        for name, t in tens.items():
            if name == "psi":
                continue
            slot, i, j = name.split("_")
            r = tl.to_ref(t)
            keep.append(r)  # the driver copies; the handle may go after set_tensor
            d.set_tensor(slot, int(i), int(j), r)
        vin = tl.to_ref(tens["psi"], 1)

    # H_eff applications within the budget
    sec = d.time_heff(il, ir, vin, 1)
    reps = 1
    if sec < a.budget / 2:
        reps = max(1, int(a.budget / sec) - 1)
        sec = d.time_heff(il, ir, vin, reps)
        reps += 1
    res["heff_s"] = sec
    res["heff_reps"] = reps
    if a.out:
        out = tl.from_ref(d.heff(il, ir, vin))
        tl.save_tensors(a.out, {"heff": out})

    if a.workload == "real" and not a.no_step:
        # environment updates of the step (the truncated-basis ones: one block
        # Hamiltonian and the block operators on each side)
        t_l, _ = _timed(lambda: d.call("left1", il))
        t_r, _ = _timed(lambda: d.call("right1", ir))
        res["env_left_s"] = t_l
        res["env_right_s"] = t_r
        if not a.no_svd:
            t_s, w = _timed(lambda: d.svd(vin, il, ir, 1.0, 0.0, cap=16384))
            res["svd_s"] = t_s
            res["svd_kept_sum_sq"] = float(np.sum(np.asarray(w)[:D] ** 2))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
