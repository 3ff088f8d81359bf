#!/usr/bin/env python3
"""H_eff*psi throughput and the mixed-vs-FP64 sweep of the two-site DMRG
inner loop on one MI355X.

`value`: H_eff*psi at the centre bond of the SU(2) J1-J2 Heisenberg cylinder
(width 6, length 24, snake ordering, 12 open bond operators across the cut)
at D = 2048 kept multiplets.  Sector dimensions follow a rotor histogram
D_J ~ (2J+1) exp(-J(J+1) / (2 sigma^2)), sigma = 1.8, fused with the
spin-1/2 site on each side; operator and wave-function entries are seeded
Gaussians.  A step is one H_eff*psi application (all 12 bond terms plus the
two block Hamiltonians) with every input resident in HBM.  The mixed path
(FP32 operands on FP32 MFMA, FP64 output) is `value`, the FP64 path is
beside it, with per-kernel HIP-event timings on the engine's stream.

Beside it: the same bond at sigma = 2.4; the sweep comparison on the 6x24
cylinder itself (iDMRG at D = 128, sweeps at 512 and 2048, then one
saturated D = 2048 sweep timed in each precision from the same state); the
H_eff of that state's own centre bond (captured environments) and its
multiplet histogram; and the reference library on the host cores in both
BLAS settings (OpenMP over bond terms; MKL-threaded), on the synthetic bond
and on the real centre step (H_eff, environment updates, truncation).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

PEAK_FP64 = 78.6    # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP32 = 157.3   # TFLOP/s, MI355X dense FP32 matrix
PEAK_BF16 = 2500.0  # TFLOP/s, MI355X dense BF16 matrix
PEAK_HBM = 8000.0   # GB/s


class HipTimer:
    """HIP events recorded on the engine's own stream."""

    def __init__(self, stream):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        self.stream = ctypes.c_void_p(stream)

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e):
        assert self.hip.hipEventRecord(e, self.stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        t = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(t), a, b) == 0
        return t.value


def time_stage(plan, timer, d_in, d_out, prec, stage, reps):
    e0, e1 = timer.event(), timer.event()
    timer.record(e0)
    for _ in range(reps):
        plan.apply_stage(d_in, d_out, prec, stage)
    timer.record(e1)
    return timer.elapsed_ms(e0, e1) / reps


def cpu_baseline(wl, budget_s=20.0):
    """The reference's hamiltonian_vector_multiplication_idmrg (dmrg_inst.cpp:283-371)
    on the same tensors, in this process: host cores, the reference's OpenMP
    over bond terms, sequential MKL."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refbind as rb
    import sadmrg_testlib as tl
    if not rb.available():
        return None
    threads = max(1, min(wl["info"]["n_op"], host_info()["usable_cpus"]))
    d, il, ir = tl.reference_heff_setup(rb, wl, threads)
    vin = tl.to_ref(wl["psi"], 1)
    t0 = time.time()
    sec = d.time_heff(il, ir, vin, 1)
    reps = 1
    if sec < budget_s / 3:
        reps = max(1, int(budget_s / sec) - 1)
        sec = d.time_heff(il, ir, vin, reps)
    out = tl.from_ref(d.heff(il, ir, vin))
    return {"seconds_per_apply": sec, "reps": reps + 1, "threads": threads, "wall": time.time() - t0,
            "output": out}


def host_info():
    """Host cores this process may use, the machine's logical CPUs and the CPU model."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"usable_cpus": usable, "logical_cpus": os.cpu_count(), "model": model}


def cpu_leg(workload, blas, threads, budget, tensors=None, out=None, step=True, svd=True, timeout=300):
    """One reference timing in a child process (oracle/ref_centre_step.py): MKL
    fixes its threading layer at a process's first BLAS call, so each BLAS
    setting runs in a process of its own.  None if the reference library is
    not built or the child fails."""
    import subprocess
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "ref_centre_step.py"), "--workload", workload,
           "--blas", blas, "--threads", str(threads), "--budget", str(budget)]
    if tensors:
        cmd += ["--tensors", tensors]
    if out:
        cmd += ["--out", out]
    if not step:
        cmd.append("--no-step")
    if not svd:
        cmd.append("--no-svd")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    except subprocess.TimeoutExpired:
        print(f"[bench] cpu leg {workload}/{blas} timed out", file=sys.stderr, flush=True)
        return None
    if r.returncode != 0:
        print(f"[bench] cpu leg {workload}/{blas} failed: {r.stderr[-400:]}", file=sys.stderr, flush=True)
        return None
    try:
        res = json.loads(r.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return None
    return None if "error" in res else res


def time_plan(plan, ctx, timer, d_in, d_out, prec, steps, warmup):
    """Wall time per application over `steps` after `warmup`, and HIP-event
    times of each stage on the engine's stream."""
    for _ in range(warmup):
        plan.apply_device(d_in, d_out, prec)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.apply_device(d_in, d_out, prec)
    ctx.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    s1 = time_stage(plan, timer, d_in, d_out, prec, 1, max(3, steps // 2))
    s2 = time_stage(plan, timer, d_in, d_out, prec, 2, max(3, steps // 2))
    st = plan.stats()
    return {"ms_per_step": wall_ms, "stage1_ms": s1, "stage2_ms": s2, "tile_cfg": [st["tile_cfg1"], st["tile_cfg2"]]}


def warm_up(sad, ctx, width, length, D, warm, j2=0.5, refine_tol=0.0):
    """SU(2) J1-J2 cylinder (width x length, snake ordering): iDMRG at warm[0],
    then one mixed sweep per further entry of warm and one at D.  The sweep
    ends at the centre bond (dmrg_inst.cpp:919-937)."""
    ns = width * length
    coup = np.zeros(ns + 1)
    coup[1], coup[2] = 1.0, j2
    hmap = sad.j1j2_cylinder(width, length)
    d = sad.DMRG(ctx, sad.SU2, 0, length, width, D)
    d.set_couplings(hmap, coup)
    d.set_precision(sad.FP32, refine=True, refine_tol=refine_tol)
    t0 = time.perf_counter()
    d.set_max_dcut(warm[0])
    d.idmrg()
    for dw in list(warm[1:]) + [D]:
        d.set_max_dcut(dw)
        d.sweep()
        print(f"[bench] warm-up sweep D={dw}: E={d.energy():.12f} ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
    d.records()
    return d, hmap, coup, time.perf_counter() - t0


def real_centre(sad, tl, ctx, d, width, length, D, hmap, coup, args, timer, save_path=None):
    """H_eff*psi at the centre bond of the warmed-up state: the plan the sweep
    builds from the driver's own environments, FP32 and FP64, the FP32 output
    error, and the bond's multiplet histogram.  With save_path, also writes
    the step's tensors for the reference CPU leg."""
    ns = width * length
    il, ir = ns // 2 - 1, ns // 2
    psi = d.input_vector(il, ir)
    _, _, ub = tl.parse_header(d.tensor("uu", il).header())
    hist = {"two_j": [int(x) for x in ub[-1][1]], "multiplets": [int(x) for x in ub[-1][2]]}
    plan = d.heff_plan(il, ir, psi)
    st = plan.stats()
    out = psi.copy()
    d_in, d_out = psi.device_ptr(), out.device_ptr()
    res = {"il": il, "ir": ir, "n": st["n"], "flops_per_apply": st["flops"], "bond_histogram": hist,
           "kept": int(sum(hist["multiplets"]))}
    for name, prec in (("fp64", sad.FP64), ("fp32", sad.FP32)):
        r = time_plan(plan, ctx, timer, d_in, d_out, prec, max(5, args.steps // 2), args.warmup)
        r["tflops"] = st["flops"] / (r["ms_per_step"] * 1e-3) / 1e12
        res[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
    o64 = tl.from_dev(plan.apply(psi, sad.FP64))
    o32 = tl.from_dev(plan.apply(psi, sad.FP32))
    res["fp32_rel_err_vs_fp64"] = tl.compare(o32, o64)[0]
    print(f"[bench] real centre bond: n={st['n']} kept={res['kept']} fp64 {res['fp64']['ms_per_step']} ms, "
          f"fp32 {res['fp32']['ms_per_step']} ms, fp32 err {res['fp32_rel_err_vs_fp64']:.2e}",
          file=sys.stderr, flush=True)
    if save_path:
        tens = {"psi": tl.from_dev(psi)}

        def put(slot, i, j=0):
            t = d.tensor(slot, i, j)
            if t.nelem() > 0:
                tens[f"{slot}_{i}_{j}"] = tl.from_dev(t)

        for i in (il, ir):
            put("uu", i)
        for i in (il - 1, il, ir, ir + 1):
            put("hh", i)
            for j in range(ns):
                put("opr", i, j)
        tl.save_tensors(save_path, tens, params=np.asarray([width, length, D, il, ir]),
                        hmap=np.asarray(hmap, np.int32), coup=np.asarray(coup))
        res["reference_output"] = o64
    del plan
    return res


# Template arguments of grouped_gemm_kernel per tile configuration (kernels.hip):
# <T, TOut, TM, TN, TK, waves per SIMD, prefetch depth>.
_CFG_ARGS = {"fp32": {0: "64, 64, 32, 2, 2", 1: "128, 128, 32, 2, 1", 2: "128, 64, 32, 2, 2", 3: "32, 32, 32, 2, 2"},
             "fp64": {0: "64, 64, 16, 2, 2", 1: "128, 128, 16, 1, 2", 2: "128, 64, 16, 2, 2", 3: "32, 32, 16, 2, 2"}}
N_TILE_CFGS = 4       # 64x64, 128x128, 128x64, 32x32
N_PLANS = 2 * N_TILE_CFGS


def gemm_kernel_name(prec, stage, plan, vec=True):
    """Kernel symbol of an H_eff stage (stage 1 writes FP32 intermediates on
    the FP32 path, stage 2 writes the FP64 output) for a launch plan id
    (tile configuration + 4 x split variant + 8 x launch kind); vec: the
    16-byte staging instantiation the H_eff plans run.  The persistent FP64
    128 x 64 kernel stages one K step ahead."""
    persist = plan >= N_PLANS
    cfg = plan % N_TILE_CFGS  # launch plan -> tile configuration
    args = _CFG_ARGS[prec][cfg]
    if persist and prec == "fp64" and cfg == 2:
        args = "128, 64, 16, 2, 1"
    tail = (", true" if vec else ", false") + (", true" if persist else ", false")
    if prec == "fp64":
        return f"sadmrg::grouped_gemm_kernel<double, double, {args}{tail}>"
    tout = "float" if stage == 1 else "double"
    return f"sadmrg::grouped_gemm_kernel<float, {tout}, {args}{tail}>"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, calibrated
    against a copy of known size) from the committed rocprofv3 counter
    summary of this workload; None if the summary does not cover it."""
    path = os.path.join(ROOT, "profiles", "grouped_gemm_heff_su2_D2048_hbm_counters.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if k is None else int(k["read_bytes"] + k["write_bytes"])


def c2_run(sad, args):
    """C2 (SU(2) open chain L=100, D=512 multiplets, FP64 throughout): the
    centre-bond H_eff*psi (n_op = 1) as one grouped launch per stage vs one
    launch per output block, beside the reference CPU path on the same
    tensors; the sector histogram of the bond and of psi."""
    import sadmrg_testlib as tl
    wl = tl.build_workload(D=512, n_op=1, sigma=1.2)
    ctx = sad.Context(0)
    cgc = sad.Cgc(1)
    psi = tl.to_dev(ctx, wl["psi"], 1)
    hl, hr = tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"])
    o1 = [tl.to_dev(ctx, t) for t in wl["o1"]]
    o2 = [tl.to_dev(ctx, t) for t in wl["o2"]]
    plan = sad.Heff(ctx, cgc, hl, hr, o1, o2, psi)
    st = plan.stats()
    out, out_u = psi.copy(), psi.copy()
    d_in = psi.device_ptr()
    timer = HipTimer(ctx.stream())
    res = {}
    for name, unb, d_out in (("grouped", False, out.device_ptr()), ("per_block", True, out_u.device_ptr())):
        for _ in range(args.warmup):
            plan.apply_device(d_in, d_out, sad.FP64, unbatched=unb)
        ctx.sync()
        e0, e1 = timer.event(), timer.event()
        timer.record(e0)
        for _ in range(args.steps):
            plan.apply_device(d_in, d_out, sad.FP64, unbatched=unb)
        timer.record(e1)
        ms = timer.elapsed_ms(e0, e1) / args.steps
        res[name] = {"ms": round(ms, 4), "tflops": round(st["flops"] / (ms * 1e-3) / 1e12, 3)}
        print(f"[bench] C2 {name}: {ms:.3f} ms {res[name]['tflops']:.2f} TFLOP/s", file=sys.stderr, flush=True)
    rel_u, _ = tl.compare(tl.from_dev(out_u), tl.from_dev(out))
    cpu = None if args.no_cpu else cpu_baseline(wl, args.cpu_budget)
    line = {"workload": "C2 centre bond: SU(2) open chain L=100, D=512 multiplets, n_op=1, FP64",
            "n": st["n"], "flops_per_apply": st["flops"], "stage_problems": [st["stage1_problems"], st["stage2_problems"]],
            "bond_histogram": {"two_j": wl["info"]["two_j_block"], "multiplets": wl["info"]["dims_block"]},
            "psi_sectors": {"two_j": wl["info"]["two_j_fused"], "dims": wl["info"]["dims_fused"]},
            "grouped": res["grouped"], "per_block": res["per_block"],
            "grouped_over_per_block": round(res["per_block"]["ms"] / res["grouped"]["ms"], 3),
            "per_block_rel_diff": rel_u}
    if cpu is not None:
        relc, _ = tl.compare(tl.from_dev(out), cpu["output"])
        line["cpu_reference"] = {"ms": round(cpu["seconds_per_apply"] * 1e3, 3), "threads": cpu["threads"],
                                 "rel_diff": relc,
                                 "grouped_over_cpu": round(cpu["seconds_per_apply"] * 1e3 / res["grouped"]["ms"], 2)}
    return line


def c2_sweeps(sad, ctx, L=100, D=512, nsweep=6):
    """C2 as whole runs: SU(2) open Heisenberg chain L=100, total spin 0,
    iDMRG and 6 sweeps at D=512 multiplets, FP64 throughout, once with the
    grouped H_eff launch and once with one launch per output block (every
    H_eff of the run, sadmrg_heff_per_block); per-sweep energies and
    truncation errors, and the centre bond's multiplet histogram."""
    import sadmrg_testlib as tl
    hmap = np.asarray(sad.open_chain(L), dtype=np.int32)
    coup = np.zeros(L + 1)
    coup[1] = 1.0
    out = {}
    for name, per_block in (("grouped", False), ("per_block", True)):
        sad.heff_per_block(per_block)
        sad.lib().sadmrg_rng_seed(42)
        d = sad.DMRG(ctx, sad.SU2, 0, 1, L, D)
        d.set_couplings(hmap, coup)
        d.set_max_dcut(D)
        ctx.sync()
        t0 = time.perf_counter()
        d.idmrg()
        d.records()
        t_idmrg = time.perf_counter() - t0
        es, trs, split = [], [], {"t_env": 0.0, "t_lanczos": 0.0, "t_svd": 0.0}
        mv = 0
        for s_ in range(nsweep):
            d.sweep()
            recs = d.records()
            es.append(d.energy())
            trs.append(max(r["trunc_err"] for r in recs))
            mv += int(sum(r["matvec64"] for r in recs))
            for k in split:
                split[k] += sum(r[k] for r in recs)
            print(f"[bench] C2 {name}: sweep {s_ + 1} E={d.energy():.12f} ({time.perf_counter() - t0:.1f} s)",
                  file=sys.stderr, flush=True)
        ctx.sync()
        secs = time.perf_counter() - t0
        _, _, ub = tl.parse_header(d.tensor("uu", L // 2 - 1).header())
        out[name] = {"seconds": round(secs, 3), "idmrg_seconds": round(t_idmrg, 3),
                     "sweep_seconds": round(secs - t_idmrg, 3), "sweep_energies": es, "max_trunc_err": trs,
                     "matvecs": mv, "split_s": {k: round(v, 3) for k, v in split.items()}}
        if name == "grouped":
            out["centre_bond_histogram"] = {"two_j": [int(x) for x in ub[-1][1]],
                                            "multiplets": [int(x) for x in ub[-1][2]]}
        del d
    sad.heff_per_block(False)
    out["workload"] = f"C2: SU(2) open chain L={L}, S=0, iDMRG + {nsweep} sweeps at D={D} multiplets, FP64"
    out["grouped_over_per_block"] = round(out["per_block"]["seconds"] / out["grouped"]["seconds"], 3)
    eg, ep = np.asarray(out["grouped"]["sweep_energies"]), np.asarray(out["per_block"]["sweep_energies"])
    # the two runs differ only in the summation order of the H_eff products
    out["rel_dE_per_sweep_grouped_vs_per_block"] = [float(x) for x in np.abs(eg - ep) / np.abs(eg)]
    return out


def c1_run(sad, args):
    """C1 (U(1) open chain L=32, m=64 kept states, 4 sweeps, FP64): the
    engine's iDMRG + 4 sweeps against the reference library's own step
    sequence on the host (same lattice, same seed); per-sweep energies and
    truncation errors side by side, wall times of both."""
    L, m, nsweep = 32, 64, 4
    hmap = np.asarray(sad.open_chain(L), dtype=np.int32)
    coup = np.zeros(L + 1)
    coup[1] = 1.0
    ctx = sad.Context(0)

    def run(device_steps, graph):
        prev = sad.lanczos_mode(device_steps, graph, reset=True)
        sad.lib().sadmrg_rng_seed(42)
        d = sad.DMRG(ctx, sad.U1, 0, 1, L, m)
        d.set_couplings(hmap, coup)
        d.set_max_dcut(m)
        ctx.sync()
        h0 = ctx.host_stats()
        t0 = time.perf_counter()
        d.idmrg()
        es, trs, t_lan = [], [], 0.0
        split = {"t_env": 0.0, "t_lanczos": 0.0, "t_svd": 0.0}
        for r in d.records():
            for k in split:
                split[k] += r[k]
        for _ in range(nsweep):
            d.records()
            d.sweep()
            recs = d.records()
            es.append(d.energy())
            trs.append(recs[-1]["trunc_err"])
            t_lan += sum(r["t_lanczos"] for r in recs)
            for r in recs:
                for k in split:
                    split[k] += r[k]
        ctx.sync()
        secs = time.perf_counter() - t0
        h1 = ctx.host_stats()
        replays = sad.lanczos_mode(prev[0], prev[1])[2]
        run.split = {k: round(v, 3) for k, v in split.items()}
        run.host = {"syncs": h1["syncs"] - h0["syncs"], "sync_s": round(h1["sync_s"] - h0["sync_s"], 3),
                    "mallocs": h1["mallocs"] - h0["mallocs"]}
        return secs, es, trs, t_lan, replays

    run(True, True)  # first use: autotuner entries, rocSOLVER and allocator warm-up
    modes = {}
    for name, dev, graph in (("host_steps", False, False), ("device_steps", True, False),
                             ("graph_steps", True, True)):
        secs, es, trs, t_lan, replays = run(dev, graph)
        modes[name] = {"seconds": round(secs, 3), "sweep_lanczos_seconds": round(t_lan, 3),
                       "graph_replays": replays, "last_energy": es[-1]}
    sad.persistent(1)
    p_s, p_e, _, p_lan, _ = run(True, True)
    sad.persistent(0)
    o_s, o_e, _, o_lan, _ = run(True, True)
    sad.persistent(2)
    gpu_s, gpu_e, gpu_tr, _, _ = run(True, True)
    line = {"workload": "C1: U(1) open chain L=32, m=64 kept states, iDMRG + 4 sweeps, FP64",
            "gpu_seconds": round(gpu_s, 3), "gpu_split_s": run.split, "gpu_host": run.host,
            "gpu_sweep_energies": gpu_e, "gpu_last_step_trunc": gpu_tr,
            "lanczos_step_modes": modes,
            "launch_kind": {"persistent": {"seconds": round(p_s, 3), "sweep_lanczos_seconds": round(p_lan, 3),
                                           "same_energies": p_e == gpu_e},
                            "one_tile_per_workgroup": {"seconds": round(o_s, 3),
                                                       "sweep_lanczos_seconds": round(o_lan, 3),
                                                       "same_energies": o_e == gpu_e}}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        import refbind as rb
        if rb.available():
            from make_golden import ref_sweep
            rb.init(physpn=1, max_angm=20, max_dcut=m, psize=1)
            r = rb.RefDMRG(0, 0, 1, L, m, 1, 0)
            r.set_couplings(hmap, coup)
            t1 = time.perf_counter()
            r.idmrg()
            cpu_e, cpu_tr = [], []
            for _ in range(nsweep):
                recs = ref_sweep(r, L, 2, None, None, d_max=m)
                cpu_e.append(recs[-1][2])
                cpu_tr.append(recs[-1][3])
            cpu_s = time.perf_counter() - t1
            line.update({"cpu_reference_seconds": round(cpu_s, 3), "cpu_threads": 1,
                         "cpu_sweep_energies": cpu_e, "cpu_last_step_trunc": cpu_tr,
                         "max_rel_dE_last_sweep": abs(gpu_e[-1] - cpu_e[-1]) / abs(cpu_e[-1])})
    return line


def c4_run(sad, args):
    """C4 (64 independent U(1) complex XXZ + Dzyaloshinskii-Moriya chains,
    L in {32, 48, 64, 96}, Delta ~ U[0.5, 1.5], D_z ~ U[0, 0.5], from seed 7,
    m = 256): the centre-bond H_eff psi of every instance in one launch per
    stage (merged block tables) against one plan per instance, complex128
    against complex64 operands, and the reference's CPU path on a sample of
    instances.  Sector structure per instance from its length; entries
    synthetic, with the instance's channel weights (1 - i D_z, Delta,
    1 + i D_z) on the bond operator."""
    import sadmrg_testlib as tl
    n_inst, m = 64, 256
    rng = np.random.default_rng(7)
    lengths = [int(x) for x in rng.choice([32, 48, 64, 96], size=n_inst)]
    delta = rng.uniform(0.5, 1.5, n_inst)
    dz = rng.uniform(0.0, 0.5, n_inst)
    wls = []
    for i, L in enumerate(lengths):
        wl = tl.build_workload_u1(m=m, length=L, seed=8 + i, cplx=1)
        o2 = wl["o2"][0]
        _, _, bonds = tl.parse_header(o2["header"])
        w = {-2: 1 - 1j * dz[i], 0: delta[i] + 0j, 2: 1 + 1j * dz[i]}
        off = 0
        for b in range(tl.nten(bonds)):
            if not o2["present"][b]:
                continue
            lab, _ = tl.block_labels(bonds, b)
            k = int(np.prod(tl.block_dims(bonds, b)))
            z = o2["data"][2 * off:2 * (off + k):2] + 1j * o2["data"][2 * off + 1:2 * (off + k):2]
            z = z * w[lab[1]]
            o2["data"][2 * off:2 * (off + k):2], o2["data"][2 * off + 1:2 * (off + k):2] = z.real, z.imag
            off += k
        wls.append(wl)
    ctx = sad.Context(0)
    plans, ins, outs = [], [], []
    for wl in wls:
        psi = tl.to_dev(ctx, wl["psi"])
        plans.append(sad.Heff(ctx, None, tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"]),
                              [tl.to_dev(ctx, wl["o1"][0])], [tl.to_dev(ctx, wl["o2"][0])], psi))
        ins.append(psi)
        outs.append(psi.copy())
    merged = sad.Heff.merge(plans)
    st = merged.stats()
    data = np.concatenate([wl["psi"]["data"] for wl in wls])
    n = data.size // 2
    vin = tl.to_dev(ctx, dict(sym=0, cplx=1, header=tl.make_header(0, [(1, [0], [n]), (-1, [0], [1])]),
                              present=np.ones(1, np.uint8), data=data))
    vout = vin.copy()
    flops = st["flops"]

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    timer = HipTimer(ctx.stream())
    f1, f2 = merged.stage_flops(1), merged.stage_flops(2)
    planar_prev = sad.heff_complex_planar()

    def kernel_path(prec):
        r = time_plan(merged, ctx, timer, vin.device_ptr(), vout.device_ptr(), prec, args.steps, args.warmup)
        iss = [merged.issued_flops(prec, 1), merged.issued_flops(prec, 2)]
        return {"ms": round(r["ms_per_step"], 4), "stage_ms": [round(r["stage1_ms"], 4), round(r["stage2_ms"], 4)],
                "tflops": round(flops / (r["ms_per_step"] * 1e-3) / 1e12, 3),
                "stage_tflops": [round(f1 / (r["stage1_ms"] * 1e-3) / 1e12, 3),
                                 round(f2 / (r["stage2_ms"] * 1e-3) / 1e12, 3)],
                "launch_plans": r["tile_cfg"],
                "mfma_fill": [round(f1 / iss[0], 4), round(f2 / iss[1], 4)]}

    res = {}
    try:
        for name, prec in (("complex128", sad.FP64), ("complex64", sad.FP32)):
            sad.heff_complex_planar(False)
            ms_m = timed(lambda: merged.apply_device(vin.device_ptr(), vout.device_ptr(), prec))
            ms_p = timed(lambda: [p.apply_device(a.device_ptr(), b.device_ptr(), prec)
                                  for p, a, b in zip(plans, ins, outs)])
            ckern = kernel_path(prec)
            sad.heff_complex_planar(True)
            planar = kernel_path(prec)
            res[name] = {"merged_ms": round(ms_m, 4), "per_instance_ms": round(ms_p, 4),
                         "merged_tflops": round(flops / (ms_m * 1e-3) / 1e12, 3),
                         "batching_speedup": round(ms_p / ms_m, 3),
                         "planar_real_tables": planar, "complex_kernel": ckern,
                         "complex_kernel_over_planar": round(planar["ms"] / ckern["ms"], 3)}
            print(f"[bench] c4 {name}: merged {ms_m:.3f} ms, per-instance {ms_p:.3f} ms, planar {planar}, "
                  f"complex kernel {ckern}", file=sys.stderr, flush=True)
    finally:
        sad.heff_complex_planar(planar_prev)
    merged.apply_device(vin.device_ptr(), vout.device_ptr(), sad.FP64)
    o64 = tl.from_dev(vout)["data"]
    per = np.concatenate([tl.from_dev(p.apply(a, sad.FP64))["data"] for p, a in zip(plans, ins)])
    merged.apply_device(vin.device_ptr(), vout.device_ptr(), sad.FP32)
    o32 = tl.from_dev(vout)["data"]
    line = {"workload": "C4: 64 U(1) complex XXZ+DM chains (L in {32,48,64,96}, seed 7), m=256, centre-bond "
                        "H_eff*psi of all instances per step",
            "instances": n_inst, "m": m, "lengths": {str(L): lengths.count(L) for L in (32, 48, 64, 96)},
            "n_total_complex": int(n), "flops_per_step": flops, "stage_problems": [st["stage1_problems"], st["stage2_problems"]],
            "launches_per_step": {"merged": 2, "per_instance": 2 * n_inst},
            **res,
            "paths": "merged/per-instance figures: the default path (the complex kernel on interleaved data); "
                     "complex_kernel: that path's stage split; planar_real_tables: complex tables lowered to real "
                     "tables on planar (re, im) operands; mfma_fill: algorithmic flops over the MFMA flops the launch plans issue (edge tiles "
                     "whole, K rounded up to the K step) per stage",
            "merged_vs_per_instance_rel_fp64": float(np.max(np.abs(o64 - per)) / np.max(np.abs(per))),
            "complex64_rel_err_vs_complex128": float(np.max(np.abs(o32 - o64)) / np.max(np.abs(o64)))}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import refbind as rb
        if rb.available():
            sample = list(range(0, n_inst, 16))
            secs = 0.0
            for i in sample:
                d, il, ir = tl.reference_heff_setup(rb, wls[i], 1)
                secs += d.time_heff(il, ir, tl.to_ref(wls[i]["psi"]), 3)
            per_inst = secs / len(sample)
            line["cpu_reference"] = {"seconds_per_instance_apply": per_inst, "threads": 1,
                                     "sample": f"instances {sample}, 3 applications each (complex128, MKL sequential)",
                                     "seconds_all_instances": per_inst * n_inst,
                                     "gpu_merged_complex64_speedup": round(per_inst * n_inst /
                                                                           (res["complex64"]["merged_ms"] * 1e-3), 1)}
    if args.c4_solves:
        line["batched_solves"] = c4_solves(sad, ctx, lengths, delta, dz, m)
    return line


def c4_solves(sad, ctx, lengths, delta, dz, m, seed=17):
    """C4 as batched ground-state solves: one U(1) complex driver per
    instance (XXZ + z-axis DM chain, its own Delta and D_z), iDMRG and one
    sweep at m kept states; the centre-bond H_eff of every instance merged
    into one plan, and all 64 Lanczos solves run together
    (sadmrg_lanczos_batched: one merged application per Krylov step) from
    the same seeded starts as 64 single complex128 solves.  The starts are
    the sweep's input vectors with 1e-3 relative Gaussian noise: warm
    starts, as a sweep step has."""
    import sadmrg_testlib as tl
    rng = np.random.default_rng(seed)
    plans, starts = [], []
    t0 = time.perf_counter()
    for i, L in enumerate(lengths):
        hmap = np.asarray(sad.open_chain(L), dtype=np.int32)
        coup = np.zeros(L + 1)
        coup[1] = 1.0
        d = sad.DMRG(ctx, sad.U1, 1, 1, L, m)
        d.set_couplings(hmap, coup)
        d.set_channel_weights(sad.xxz_dm_weights(L + 1, {1: (1.0, float(delta[i]), float(dz[i]))}))
        d.set_max_dcut(m)
        d.idmrg()
        d.sweep()
        d.records()
        il, ir = L // 2 - 1, L // 2
        psi = d.input_vector(il, ir)
        h = tl.from_dev(psi)
        h["data"] = h["data"] + 1e-3 * np.abs(h["data"]).max() * rng.standard_normal(h["data"].size)
        starts.append(h)
        plans.append(d.heff_plan(il, ir, psi))
        del d
        print(f"[bench] c4 solves: driver {i} (L={L}) ready at {time.perf_counter() - t0:.1f} s", file=sys.stderr,
              flush=True)
    t_prep = time.perf_counter() - t0
    print(f"[bench] c4 solves: {len(plans)} drivers prepared in {t_prep:.1f} s", file=sys.stderr, flush=True)
    merged = sad.Heff.merge(plans)
    data = np.concatenate([s["data"] for s in starts])
    n = data.size // 2
    hdr = tl.make_header(0, [(1, [0], [n]), (-1, [0], [1])])

    def batched(prec, refine=True):
        v = tl.to_dev(ctx, dict(sym=0, cplx=1, header=hdr, present=np.ones(1, np.uint8), data=data))
        opts = dict(mlanc=40, precision=prec)
        if refine:
            opts.update(refine=1, refine_tol=1e-8, refine_krylov=8, max_restarts=16)
        ctx.sync()
        t1 = time.perf_counter()
        e, info = merged.lanczos_batched(v.device_ptr(), len(plans), **opts)
        ctx.sync()
        return e, info, time.perf_counter() - t1

    batched(sad.FP32)  # first use: tile plans and allocator
    e32, i32, t32 = batched(sad.FP32)
    e64, i64, t64 = batched(sad.FP64)            # the same refinement in complex128: the reference
    e64p, i64p, t64p = batched(sad.FP64, False)  # lanczos1's own stop only
    ctx.sync()
    t1 = time.perf_counter()
    es = []
    for p, s in zip(plans, starts):
        v = tl.to_dev(ctx, s)
        e, _ = p.lanczos(v, mlanc=40, precision=sad.FP64)
        es.append(e)
    ctx.sync()
    t_single = time.perf_counter() - t1
    es = np.asarray(es)
    rel32 = np.abs(e32 - es) / np.abs(es)
    rel64 = np.abs(e64p - es) / np.abs(es)
    rel3264 = np.abs(e32 - e64) / np.abs(e64)
    return {"drivers": f"U(1) complex XXZ+DM chains, iDMRG + one sweep at m={m} each ({t_prep:.1f} s)",
            "n_total_complex": int(n), "solves": len(plans),
            "batched_mixed": {"seconds": round(t32, 4), "solves_per_s": round(len(plans) / t32, 1),
                              "merged_matvecs": [i32["matvec32"], i32["matvec64"]],
                              "max_restarts": int(i32["restarts"].max()),
                              "max_residual_rel": float(np.max(i32["residual64"] / np.abs(e32)))},
            "batched_complex128_refined": {"seconds": round(t64, 4), "solves_per_s": round(len(plans) / t64, 1),
                                           "merged_matvecs": i64["matvec64"],
                                           "max_restarts": int(i64["restarts"].max())},
            "batched_complex128": {"seconds": round(t64p, 4), "solves_per_s": round(len(plans) / t64p, 1),
                                   "merged_matvecs": i64p["matvec64"]},
            "single_complex128": {"seconds": round(t_single, 4), "solves_per_s": round(len(plans) / t_single, 1)},
            "batched_mixed_over_single": round(t_single / t32, 2),
            "max_rel_dE_mixed_vs_complex128_refined": float(rel3264.max()),
            "max_rel_dE_batched_vs_single_complex128": float(rel64.max()),
            "max_rel_dE_mixed_vs_single_complex128": float(rel32.max()),
            "instances_not_refined_to_tol": int(np.sum(i32["residual64"] > 1e-8 * np.abs(e32) * (1 + 1e-6))),
            "note": "refined solves: |r| <= 1e-8 |E| in FP64 after complex64 (mixed) or complex128 Krylov restarts; "
                    "the single and unrefined batched solves stop at lanczos1's test (|<v,Hv>|/|Hv| within 1e-8 of "
                    "1), whose energy error bounds their difference from the refined ones"}


def env_run(sad, args):
    """Environment updates at the C3 centre bond (SU(2), block bond of D
    multiplets from the rotor histogram, spin-1/2 site), FP64: the
    enlarged-basis operator_transformation (identity fusion of block and
    site: a block scatter, memory bound) and the truncated-basis one (a
    random site tensor keeping half of each fused sector: grouped GEMMs).
    Milliseconds per update from the host (table building and
    synchronisation included) and algorithmic bytes of the scatter."""
    import sadmrg_testlib as tl
    ctx = sad.Context(0)
    rng = np.random.default_rng(5)
    js, ds = tl.rotor_dims(args.D, 1.8)
    b = (1, js, ds)
    site = (1, [1], [1])
    fj, fd = tl.fused(js, ds)
    op = tl.to_dev(ctx, tl.random_tensor(rng, 1, 0, 0, [(-1, js, ds), (-1, [2], [1]), b]))
    u1 = sad.fuse(ctx, sad.SU2, 0, b, site)
    uu = tl.to_dev(ctx, tl.random_tensor(rng, 1, 0, 0, [b, site, (-1, fj, [max(1, d // 2) for d in fd])]))

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3 / args.steps, out

    cgc = sad.Cgc(1)
    ms_s, out_s = timed(lambda: sad.env("operator_transformation", u1, u1, op, flag=0, cgc=cgc))
    ms_t, out_t = timed(lambda: sad.env("operator_transformation", uu, uu, op, flag=0, cgc=cgc))
    op_b, out_b = op.nelem() * 8, out_s.nelem() * 8
    return {"workload": f"environment updates at the C3 centre bond, SU(2) block bond D={args.D} multiplets, FP64",
            "enlarged_basis_scatter": {"ms": round(ms_s, 4), "operator_bytes": op_b, "output_bytes": out_b,
                                       "algorithmic_GBps": round((op_b + out_b) / (ms_s * 1e-3) / 1e9, 1),
                                       "note": "host time per update: table build, scatter launch, synchronisation"},
            "truncated_basis_gemm": {"ms": round(ms_t, 4), "output_bytes": out_t.nelem() * 8}}


def sweep_compare(sad, ctx, d, width, length, D, warm, j2=0.5, refine_tol=0.0, refine_krylov=6, check_every=(4,),
                  bf16=False, after_fp64=None, fp64_every=()):
    """Mixed vs FP64 sweep on the warmed-up SU(2) J1-J2 cylinder (width x
    length, snake ordering; see warm_up): the state is cloned on the device
    and one sweep is timed in each precision from the identical start.  The
    FP64 sweep keeps the reference's Lanczos schedule (check every 4th step);
    the mixed sweep checks every check_every[0]-th step, further entries are
    timed as variants.  after_fp64(c), if given, runs on the FP64 clone after
    its sweep (the saturated state at the centre bond), before the clone is
    released: at most two states are resident at a time."""
    ns = width * length
    out = {}
    after = None
    memory = {}
    legs = [("fp64", sad.FP64, 4, 1), ("mixed", sad.FP32, check_every[0], 1)]
    legs += [(f"mixed_check{k}", sad.FP32, k, 1) for k in check_every[1:]]
    # FP64 energy at the sweep's last step only; the other steps stop on the
    # FP32 recurrence residual and report FP32 Ritz values
    legs += [(f"mixed_fp64_every{k}", sad.FP32, check_every[0], k) for k in fp64_every]
    if bf16:  # ablation: BF16 Krylov matvecs, same FP64 checks and refinement
        legs.append(("bf16", sad.BF16, check_every[0], 1))
    for name, prec, every, e64 in legs:
        c = d.clone()
        c.set_precision(prec, refine=(prec != sad.FP64), refine_tol=refine_tol)
        c.set_refinement(refine_krylov, 6)
        c.set_check_interval(every)
        c.set_fp64_energy_every(e64)
        ctx.sync()
        sad.svd_stats(reset=True)
        sad.lanczos_mode(reset=True)
        t1 = time.perf_counter()
        c.sweep()
        ctx.sync()
        wall = time.perf_counter() - t1
        recs = c.records()
        out[name] = {"seconds": wall, "energy": c.energy(), "recs": recs, "svd": sad.svd_stats(reset=True),
                     "graph_replays": sad.lanczos_mode()[2]}
        print(f"[bench] {name} sweep: {wall:.2f} s, E={c.energy():.12f}, steps={len(recs)}",
              file=sys.stderr, flush=True)
        memory[name] = ctx.memory()
        if name == "fp64" and after_fp64 is not None:
            after = after_fp64(c)
        del c
    e64 = np.asarray([r["energy"] for r in out["fp64"]["recs"]])
    emx = np.asarray([r["energy"] for r in out["mixed"]["recs"]])
    rel = np.abs(emx - e64) / np.abs(e64)
    # the FP32 Krylov Ritz value each mixed step had before its FP64 check
    r32 = np.asarray([r["ritz32"] for r in out["mixed"]["recs"]])
    rel32 = np.abs(r32 - emx) / np.abs(emx)

    def variant(name):
        recs = out[name]["recs"]
        e = np.asarray([r["energy"] for r in recs])
        f64 = np.asarray([r["energy64"] for r in recs]) > 0
        rv = np.abs(e - e64) / np.abs(e64)
        return {"sweep_s": round(out[name]["seconds"], 3),
                "speedup": round(out["fp64"]["seconds"] / out[name]["seconds"], 3),
                "final_rel_dE": float(abs(out[name]["energy"] - out["fp64"]["energy"]) / abs(out["fp64"]["energy"])),
                "max_rel_dE_per_step": float(rv.max()),
                "fp64_energy_steps": int(f64.sum()),
                "max_rel_dE_fp64_steps": float(rv[f64].max()) if f64.any() else None,
                "split_s": {k: round(tsum(name, k), 3) for k in ("t_env", "t_lanczos", "t_svd")},
                "matvecs": [int(tsum(name, "matvec32")), int(tsum(name, "matvec64"))]}

    def tsum(name, key):
        return float(sum(r[key] for r in out[name]["recs"]))

    def centre(name):
        """The sweep's last step at the centre bond (ns/2 - 1, ns/2)."""
        rec = [r for r in out[name]["recs"] if int(r["il"]) == ns // 2 - 1 and int(r["ir"]) == ns // 2]
        if not rec:
            return None
        r = rec[-1]
        return {k: (round(float(r[k]), 5) if k.startswith("t_") else int(r[k]))
                for k in ("t_env", "t_lanczos", "t_svd", "matvec64", "matvec32", "kept")}

    return after, {
        "device_memory_bytes": {k: {"in_use": v[0], "cached": v[1]} for k, v in memory.items()},
        "lattice": f"SU(2) J1-J2 cylinder {width}x{length} (J2/J1={j2}), snake ordering, {ns} sites",
        "D": D, "warm_up": list(warm) + [D], "refine_tol": refine_tol or 1e-7, "refine_krylov": refine_krylov,
        "fp64_sweep_s": round(out["fp64"]["seconds"], 3), "mixed_sweep_s": round(out["mixed"]["seconds"], 3),
        "speedup": round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3),
        "energy_fp64": out["fp64"]["energy"], "energy_mixed": out["mixed"]["energy"],
        "max_rel_dE_per_step": float(rel.max()), "final_rel_dE": float(rel[-1]),
        "steps": len(e64), "max_kept": int(max(r["kept"] for r in out["fp64"]["recs"])),
        "fp64_split_s": {k: round(tsum("fp64", k), 3) for k in ("t_env", "t_lanczos", "t_svd")},
        "mixed_split_s": {k: round(tsum("mixed", k), 3) for k in ("t_env", "t_lanczos", "t_svd")},
        "matvecs_fp64": int(tsum("fp64", "matvec64")),
        "matvecs_mixed": [int(tsum("mixed", "matvec32")), int(tsum("mixed", "matvec64"))],
        "mixed_check_every": check_every[0],
        "centre_step": {"fp64": centre("fp64"), "mixed": centre("mixed")},
        "lanczos_graph_replays": {"fp64": out["fp64"]["graph_replays"], "mixed": out["mixed"]["graph_replays"]},
        "truncation": {"fp64": out["fp64"]["svd"], "mixed": out["mixed"]["svd"]},
        "ritz32_vs_fp64_rayleigh": {"max_rel": float(rel32.max()), "median_rel": float(np.median(rel32)),
                                    "note": "per mixed step: the FP32 Krylov Ritz value before the FP64 check "
                                            "against that step's FP64 Rayleigh quotient"},
        "variants": {name: variant(name) for name, _, _, _ in legs[2:]},
    }


def gradient_run(sad, ctx, L, D, warm, seed=11, fd_bonds=(0, None, -1), h=1e-5, sweeps=3):
    """dE/dJ_i for every bond of an open SU(2) chain with random couplings
    J_i ~ U[0.5, 1.5] (seed), at D kept multiplets: adjoint gradient timed,
    checked by Euler's identity sum_i J_i dE/dJ_i = E and by central finite
    differences on a few bonds (each side re-converged from the state)."""
    rng = np.random.default_rng(seed)
    J = rng.uniform(0.5, 1.5, L - 1)
    hmap = np.zeros((L, L), dtype=np.int32)
    coup = np.zeros(L + 1)
    for i in range(L - 1):
        hmap[i, i + 1] = hmap[i + 1, i] = i + 1
        coup[i + 1] = J[i]
    d = sad.DMRG(ctx, sad.SU2, 0, 1, L, D)
    d.set_couplings(hmap, coup)
    d.set_precision(sad.FP64)
    t0 = time.perf_counter()
    d.set_max_dcut(warm[0])
    d.idmrg()
    for dw in list(warm[1:]) + [D] * sweeps:
        d.set_max_dcut(dw)
        d.sweep()
        print(f"[bench] gradient chain L={L}: sweep D={dw} E={d.energy():.12f} ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
    d.records()
    ctx.sync()
    t1 = time.perf_counter()
    g, e = d.coupling_gradient()
    ctx.sync()
    t_grad = time.perf_counter() - t1
    euler = float(np.dot(coup, g))
    picks = sorted({(L - 1) // 2 if b is None else (b % (L - 1)) for b in fd_bonds})
    fd = []
    for b in picks:
        vals = []
        for sgn in (1.0, -1.0):
            c = d.clone()
            cp = coup.copy()
            cp[b + 1] += sgn * h
            c.set_couplings(hmap, cp)
            c.prepare_sweep()
            for _ in range(2):
                c.sweep()
            vals.append(c.energy())
            c.records()
            del c
        fdv = (vals[0] - vals[1]) / (2 * h)
        fd.append({"bond": b, "adjoint": float(g[b + 1]), "finite_difference": fdv,
                   "abs_diff": abs(fdv - float(g[b + 1]))})
        print(f"[bench] bond {b}: adjoint {g[b + 1]:.10f} fd {fdv:.10f}", file=sys.stderr, flush=True)
    return {"chain": f"SU(2) open chain L={L}, J_i ~ U[0.5,1.5] seed {seed}", "D": D, "bonds": L - 1,
            "energy": e, "sweep_energy": d.energy(), "gradient_seconds": round(t_grad, 4),
            "euler_rel_err": abs(euler - e) / abs(e), "finite_difference": fd,
            "dE_dJ": [float(x) for x in g[1:L]]}


def cpu_summary(cpu, flops, real, sweep):
    """`cpu_baseline` of the JSON line: the faster BLAS setting of the
    reference's H_eff on the synthetic bond as `value`, both settings, the
    host, and the centre-step estimate of a reference sweep step on the real
    state: (FP64 matvecs of the step) x H_eff + both truncated-basis
    environment updates + truncation, each at its faster setting."""
    if cpu is None or not cpu["synthetic"]:
        return None
    syn = cpu["synthetic"]
    best = min(syn, key=lambda k: syn[k]["heff_s"])
    b = syn[best]
    out = {"value": round(flops / b["heff_s"] / 1e12, 5), "unit": "TFLOP/s", "cores": b["threads"],
           "kind": "reference",
           "sample": f"{b['heff_reps']} applications of DMRG::hamiltonian_vector_multiplication_idmrg on the "
                     f"synthetic centre bond ({b['heff_s']:.3f} s each, FP64, "
                     + ("sequential MKL + the reference's OpenMP over bond terms" if best == "omp"
                        else "MKL-threaded BLAS, bond terms in sequence") + ")",
           "host": cpu["host"],
           "settings": {k: {"threads": v["threads"], "heff_s": round(v["heff_s"], 4),
                            "tflops": round(flops / v["heff_s"] / 1e12, 5)} for k, v in syn.items()},
           "parity_rel_vs_gpu_fp64": syn.get("omp", {}).get("parity_rel_vs_gpu_fp64")}
    rr = cpu.get("real") or {}
    if real is not None and rr and sweep is not None and sweep.get("centre_step", {}).get("fp64"):
        def fastest(key):
            vals = [v[key] for v in rr.values() if key in v]
            return min(vals) if vals else None
        heff, el, er, sv = fastest("heff_s"), fastest("env_left_s"), fastest("env_right_s"), fastest("svd_s")
        c64, cmx = sweep["centre_step"]["fp64"], sweep["centre_step"]["mixed"]
        mv = c64["matvec64"]
        est = {"heff_s": heff, "env_left_s": el, "env_right_s": er, "svd_s": sv, "fp64_matvecs": mv,
               "settings": {k: {kk: (round(vv, 4) if isinstance(vv, float) and not kk.startswith("parity") else vv)
                                for kk, vv in v.items()} for k, v in rr.items()},
               "heff_tflops": round(real["flops_per_apply"] / heff / 1e12, 5) if heff else None,
               "parity_rel_vs_gpu_fp64": rr.get("omp", {}).get("parity_rel_vs_gpu_fp64")}
        if None not in (heff, el, er, sv):
            step = mv * heff + el + er + sv
            g64 = c64["t_env"] + c64["t_lanczos"] + c64["t_svd"]
            gmx = cmx["t_env"] + cmx["t_lanczos"] + cmx["t_svd"] if cmx else None
            est.update({"reference_step_s": round(step, 3), "gpu_fp64_step_s": round(g64, 4),
                        "gpu_mixed_step_s": round(gmx, 4) if gmx else None,
                        "speedup_fp64": round(step / g64, 1), "speedup_mixed": round(step / gmx, 1) if gmx else None,
                        "note": "the reference's step estimated from its own H_eff, environment updates "
                                "(prepare_site_operator_from_left/right) and truncation (BlockTensor::svd) timed "
                                "on the real centre-bond tensors; the enlarged-basis update and the input-vector "
                                "transformation are not counted, so the estimate is low"})
        out["centre_step"] = est
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--D", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device tensor 5 times (HBM counter calibration) and exit")
    ap.add_argument("--refine-tol", type=float, default=3e-6,
                    help="FP64 residual tolerance of the mixed path (relative to |E|; 0: library default)")
    ap.add_argument("--refine-krylov", type=int, default=3, help="FP32 Krylov steps per refinement restart")
    ap.add_argument("--bf16-sweep", action="store_true", help="also time the BF16 ablation sweep")
    ap.add_argument("--sigma-variant", type=float, default=2.4,
                    help="also time the synthetic bond with this rotor-histogram width (0: skip)")
    ap.add_argument("--no-split-ablation", action="store_true",
                    help="skip the uniform-split ablation (its launches share kernel names with the main legs)")
    ap.add_argument("--lanczos-graph", default="on", choices=["on", "off"],
                    help="replay Lanczos steps from HIP graphs (off: the same kernels launched from the host, "
                         "for tools that do not follow graph launches)")
    ap.add_argument("--check-every", default="2",
                    help="mixed sweep: Lanczos convergence check interval (comma list: first is reported, "
                         "the others are timed as variants)")
    ap.add_argument("--gradient", default="none", metavar="L:D:WARM",
                    help="only run the coupling-gradient workload (e.g. 48:1024:128,256,512) and exit")
    ap.add_argument("--workload", default="c3", choices=["c3", "c2", "c1", "c4", "env"],
                    help="c3: the metric (default); c2: L=100 D=512 grouped vs per-block launch; "
                         "c1: U(1) L=32 m=64 four sweeps vs the reference; "
                         "c4: 64 U(1) complex chains at m=256, all centre bonds in one launch; "
                         "env: environment updates at the C3 centre bond")
    ap.add_argument("--extras", default="c2,c5",
                    help="default workload: further configurations reported in the same line (c2: whole runs, "
                         "grouped vs per-block; c5: the L=48 D=1024 gradient); '' for none")
    ap.add_argument("--c4-solves", type=int, default=1,
                    help="c4 workload: also run the 64 batched ground-state solves on real drivers (0: skip)")
    ap.add_argument("--fp64-energy-every", default="0",
                    help="mixed sweep variants computing the FP64 energy at every k-th step and the last "
                         "(0: the last step only); comma list, '' for none")
    ap.add_argument("--sweep", default="6x24:2048:128,512",
                    help="WIDTHxLENGTH:D:warm-up Ds for the mixed vs FP64 sweep comparison ('none' skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the engine has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import sadmrg_testlib as tl
    sad = tl.load_pkg()
    if args.lanczos_graph == "off":
        sad.lanczos_mode(graph=False)
    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 8
        ctx = sad.Context(0)
        t = tl.to_dev(ctx, dict(sym=0, cplx=0, header=tl.make_header(0, [(1, [0], [n]), (-1, [0], [1])]),
                                present=np.ones(1, np.uint8), data=np.ones(n)))
        for _ in range(5):
            t.copy()
        ctx.sync()
        print(json.dumps({"calibration": "device-to-device tensor copy", "bytes_read": n * 8,
                          "bytes_written": n * 8, "copies": 5}))
        return
    if args.workload == "c2":
        print(json.dumps(c2_run(sad, args)))
        return
    if args.workload == "c1":
        print(json.dumps(c1_run(sad, args)))
        return
    if args.workload == "c4":
        print(json.dumps(c4_run(sad, args)))
        return
    if args.workload == "env":
        print(json.dumps(env_run(sad, args)))
        return
    if args.gradient != "none":
        Ls, Ds, Ws = args.gradient.split(":")
        ctx = sad.Context(0)
        out = gradient_run(sad, ctx, int(Ls), int(Ds), [int(x) for x in Ws.split(",")])
        print(json.dumps({"workload": "coupling gradient", **out}))
        return
    print(f"[bench] building C3 workload D={args.D}", file=sys.stderr, flush=True)
    wl = tl.build_workload(D=args.D)
    ctx = sad.Context(0)
    cgc = sad.Cgc(1)
    psi = tl.to_dev(ctx, wl["psi"], 1)
    hl, hr = tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"])
    o1 = [tl.to_dev(ctx, t) for t in wl["o1"]]
    o2 = [tl.to_dev(ctx, t) for t in wl["o2"]]
    plan = sad.Heff(ctx, cgc, hl, hr, o1, o2, psi)
    st = plan.stats()
    out = psi.copy()
    d_in, d_out = psi.device_ptr(), out.device_ptr()
    timer = HipTimer(ctx.stream())
    flops = st["flops"]
    f1, f2 = plan.stage_flops(1), plan.stage_flops(2)
    print(f"[bench] n={st['n']} flops/apply={flops:.4e} tiles={st['tiles1']}+{st['tiles2']}", file=sys.stderr,
          flush=True)

    res = {}
    # persistent launch mode: 2 the autotuned default, 0 one workgroup per
    # tile everywhere, 1 persistent everywhere (same tuned tile plans)
    legs = [("fp64", sad.FP64, True, 2), ("fp32", sad.FP32, True, 2), ("bf16", sad.BF16, True, 2),
            ("fp64_elem", sad.FP64, False, 2), ("fp32_elem", sad.FP32, False, 2),
            ("fp64_one_tile", sad.FP64, True, 0), ("fp32_one_tile", sad.FP32, True, 0),
            ("fp64_persistent", sad.FP64, True, 1), ("fp32_persistent", sad.FP32, True, 1)]
    for name, prec, vec, pers in legs:
        sad.vec_staging(vec)
        sad.persistent(pers)
        for _ in range(args.warmup):
            plan.apply_device(d_in, d_out, prec)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.apply_device(d_in, d_out, prec)
        ctx.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        s1 = time_stage(plan, timer, d_in, d_out, prec, 1, max(3, args.steps // 2))
        s2 = time_stage(plan, timer, d_in, d_out, prec, 2, max(3, args.steps // 2))
        stp = plan.stats()
        res[name] = {"ms_per_step": wall_ms, "tflops": flops / (wall_ms * 1e-3) / 1e12,
                     "stage1_ms": s1, "stage2_ms": s2,
                     "stage1_tflops": f1 / (s1 * 1e-3) / 1e12, "stage2_tflops": f2 / (s2 * 1e-3) / 1e12,
                     "tile_cfg": [stp["tile_cfg1"], stp["tile_cfg2"]]}
        print(f"[bench] {name}: {wall_ms:.3f} ms/apply {res[name]['tflops']:.2f} TFLOP/s "
              f"(stage1 {s1:.3f} ms, stage2 {s2:.3f} ms)", file=sys.stderr, flush=True)
    sad.vec_staging(True)
    sad.persistent(2)

    # ablation: the uniform split policy (every stage-2 problem cut into the
    # same number of reduction parts), on a plan built and tuned under it
    uniform = None
    if not args.no_split_ablation:
        sad.balanced_split(False)
        sad.autotune(clear=True)
        uplan = sad.Heff(ctx, cgc, hl, hr, o1, o2, psi)
        uniform = {}
        for name, prec in (("fp32", sad.FP32), ("fp64", sad.FP64)):
            for _ in range(args.warmup):
                uplan.apply_device(d_in, d_out, prec)
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                uplan.apply_device(d_in, d_out, prec)
            ctx.sync()
            uniform[name + "_ms"] = round((time.perf_counter() - t0) * 1e3 / args.steps, 4)
            uniform[name + "_stage2_ms"] = round(time_stage(uplan, timer, d_in, d_out, prec, 2, max(3, args.steps // 2)), 4)
        del uplan
        sad.balanced_split(True)
        sad.autotune(clear=True)
        uniform["note"] = ("stage-2 reductions cut into the same number of parts per problem instead of parts of equal "
                           "depth sized to the resident workgroups")
        print(f"[bench] uniform split: {uniform}", file=sys.stderr, flush=True)

    # accuracy of the mixed path against FP64 on the same input
    o64 = tl.from_dev(plan.apply(psi, sad.FP64))
    o32 = tl.from_dev(plan.apply(psi, sad.FP32))
    rel32, _ = tl.compare(o32, o64)
    o16 = tl.from_dev(plan.apply(psi, sad.BF16))
    rel16, _ = tl.compare(o16, o64)

    # the sector histogram's width: the same bond with sigma = 2.4
    sig24 = None
    if args.sigma_variant > 0:
        wl2 = tl.build_workload(D=args.D, sigma=args.sigma_variant)
        psi2 = tl.to_dev(ctx, wl2["psi"], 1)
        plan2 = sad.Heff(ctx, cgc, tl.to_dev(ctx, wl2["hl"]), tl.to_dev(ctx, wl2["hr"]),
                         [tl.to_dev(ctx, t) for t in wl2["o1"]], [tl.to_dev(ctx, t) for t in wl2["o2"]], psi2)
        out2 = psi2.copy()
        st2 = plan2.stats()
        sig24 = {"sigma": args.sigma_variant, "n": st2["n"], "flops_per_apply": st2["flops"],
                 "sectors": wl2["info"]["two_j_fused"], "sector_dims": wl2["info"]["dims_fused"]}
        for name, prec in (("fp64", sad.FP64), ("fp32", sad.FP32)):
            r = time_plan(plan2, ctx, timer, psi2.device_ptr(), out2.device_ptr(), prec, max(5, args.steps // 2),
                          args.warmup)
            r["tflops"] = st2["flops"] / (r["ms_per_step"] * 1e-3) / 1e12
            sig24[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        print(f"[bench] sigma={args.sigma_variant}: fp32 {sig24['fp32']['tflops']:.2f} TF/s, "
              f"fp64 {sig24['fp64']['tflops']:.2f} TF/s", file=sys.stderr, flush=True)
        del plan2, psi2, out2, wl2

    # the sweep comparison on the metric's lattice, with the H_eff of its own
    # centre bond (captured environments) first
    sweep, real = None, None
    tmpdir = None
    if args.sweep != "none":
        import tempfile
        lat, dd, warm = args.sweep.split(":")
        w_, l_ = (int(x) for x in lat.split("x"))
        warm_d = [int(x) for x in warm.split(",")]
        print(f"[bench] sweep comparison {lat} D={dd}", file=sys.stderr, flush=True)
        dr, hmap, coup, t_warm = warm_up(sad, ctx, w_, l_, int(dd), warm_d, refine_tol=args.refine_tol)
        real_path = None
        if not args.no_cpu:
            tmpdir = tempfile.mkdtemp(prefix="sadmrg_bench_")
            real_path = os.path.join(tmpdir, "centre.npz")
        real, sweep = sweep_compare(
            sad, ctx, dr, w_, l_, int(dd), warm_d, refine_tol=args.refine_tol, refine_krylov=args.refine_krylov,
            check_every=tuple(int(x) for x in args.check_every.split(",")), bf16=args.bf16_sweep,
            fp64_every=tuple(int(x) for x in args.fp64_energy_every.split(",") if x != ""),
            after_fp64=lambda c: real_centre(sad, tl, ctx, c, w_, l_, int(dd), hmap, coup, args, timer, real_path))
        sweep["warm_up_s"] = round(t_warm, 2)
        del dr

    # further configurations in the same line: C2 whole runs (grouped vs
    # per-block launches, plus the centre-bond H_eff against the reference
    # on the host) and the C5 gradient with a finite-difference check
    extras = {}
    want = [x for x in args.extras.split(",") if x]
    if "c2" in want:
        print("[bench] C2 runs", file=sys.stderr, flush=True)
        extras["c2"] = c2_sweeps(sad, ctx)
        extras["c2"]["centre_bond_heff"] = c2_run(sad, args)
    if "c5" in want:
        print("[bench] C5 gradient", file=sys.stderr, flush=True)
        extras["c5"] = {"workload": "C5: coupling gradient, SU(2) open chain L=48, D=1024",
                        **gradient_run(sad, ctx, 48, 1024, [256, 512], fd_bonds=(0, None, -1))}

    # the reference on the host cores: both BLAS settings, each in a child
    # process; the synthetic bond for `value`, the real centre step for the
    # sweep-step estimate
    cpu = None
    host = host_info()
    if not args.no_cpu:
        import tempfile
        tmpdir = tmpdir or tempfile.mkdtemp(prefix="sadmrg_bench_")
        ncpu = host["usable_cpus"]
        n_terms = wl["info"]["n_op"]
        cpu = {"host": host, "synthetic": {}, "real": {}}
        ref_out = os.path.join(tmpdir, "ref_heff.npz")
        for blas, threads in (("omp", max(1, min(n_terms, ncpu))), ("mkl", ncpu)):
            print(f"[bench] reference CPU, synthetic bond, {blas} x {threads}", file=sys.stderr, flush=True)
            r = cpu_leg("synthetic", blas, threads, args.cpu_budget / 2, out=ref_out if blas == "omp" else None)
            if r is not None:
                cpu["synthetic"][blas] = r
                if blas == "omp" and os.path.exists(ref_out):
                    tens, _ = tl.load_tensors(ref_out)
                    r["parity_rel_vs_gpu_fp64"] = tl.compare(o64, tens["heff"])[0]
        if real is not None and real_path is not None:
            ref_out2 = os.path.join(tmpdir, "ref_heff_real.npz")
            for blas, threads in (("omp", ncpu), ("mkl", ncpu)):
                print(f"[bench] reference CPU, real centre step, {blas} x {threads}", file=sys.stderr, flush=True)
                r = cpu_leg("real", blas, threads, args.cpu_budget / 2, tensors=real_path,
                            out=ref_out2 if blas == "omp" else None, svd=(blas == "mkl"))
                if r is not None:
                    cpu["real"][blas] = r
                    if blas == "omp" and os.path.exists(ref_out2):
                        tens, _ = tl.load_tensors(ref_out2)
                        r["parity_rel_vs_gpu_fp64"] = tl.compare(real["reference_output"], tens["heff"])[0]
        for k, v in cpu.items():
            if k != "host":
                print(f"[bench] cpu {k}: {json.dumps(v)}", file=sys.stderr, flush=True)
    if real is not None:
        real.pop("reference_output", None)
    if tmpdir:
        import shutil
        shutil.rmtree(tmpdir, ignore_errors=True)

    dom = "stage2" if res["fp32"]["stage2_ms"] >= res["fp32"]["stage1_ms"] else "stage1"
    dom_kernel = gemm_kernel_name("fp32", 2 if dom == "stage2" else 1, res["fp32"]["tile_cfg"][1 if dom == "stage2" else 0])
    traffic = measured_traffic(dom_kernel)
    dom_flops = f2 if dom == "stage2" else f1
    dom_ms = res["fp32"][dom + "_ms"]
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    line = {
        "metric": "H_eff\u00b7\u03c8 TFLOP/s + sweep time, SU(2) D=2048: mixed vs FP64 GPU at |\u0394E|\u22641e-9",
        "value": round(res["fp32"]["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["fp32"]["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 operands, fp32 MFMA accumulate, fp64 output",
        "data": "synthetic",
        "config": {"workload": f"SU(2) J1-J2 6x24 cylinder centre bond, D={args.D} multiplets, n_op=12, "
                               f"N={st['n']} reduced elements, one H_eff*psi per step",
                   "flops_per_step": flops, "sectors": wl["info"]["two_j_fused"],
                   "sector_dims": wl["info"]["dims_fused"]},
        "fp64": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res["fp64"].items()},
        "fp32": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res["fp32"].items()},
        "tile_configs": "launch plan = tile configuration (0: 64x64, 1: 128x128, 2: 128x64, 3: 32x32 output tiles) "
                        "+ 4 x split variant (0: about 3, 1: about 8 reduction parts per resident workgroup) + 8 x "
                        "launch kind (0: one workgroup per tile, 1: persistent grid taking tiles from a device "
                        "ticket counter), autotuned per stage",
        "mixed_vs_fp64_speedup": round(res["fp64"]["ms_per_step"] / res["fp32"]["ms_per_step"], 3),
        "element_staging": {"fp32_ms": round(res["fp32_elem"]["ms_per_step"], 4),
                            "fp64_ms": round(res["fp64_elem"]["ms_per_step"], 4),
                            "fp32_stage_ms": [round(res["fp32_elem"]["stage1_ms"], 4),
                                              round(res["fp32_elem"]["stage2_ms"], 4)],
                            "note": "the same kernels staging one element per lane instead of 16-byte vectors"},
        "uniform_split": uniform,
        "launch_kind": {k: {"one_tile_per_workgroup_ms": round(res[f"{k}_one_tile"]["ms_per_step"], 4),
                            "one_tile_stage_ms": [round(res[f"{k}_one_tile"]["stage1_ms"], 4),
                                                  round(res[f"{k}_one_tile"]["stage2_ms"], 4)],
                            "persistent_ms": round(res[f"{k}_persistent"]["ms_per_step"], 4),
                            "persistent_stage_ms": [round(res[f"{k}_persistent"]["stage1_ms"], 4),
                                                    round(res[f"{k}_persistent"]["stage2_ms"], 4)],
                            "autotuned_ms": round(res[k]["ms_per_step"], 4)}
                        for k in ("fp64", "fp32")},
        "fp32_rel_err_vs_fp64": rel32,
        "bf16_ablation": {"ms_per_step": round(res["bf16"]["ms_per_step"], 4),
                          "tflops": round(res["bf16"]["tflops"], 3),
                          "stage_ms": [round(res["bf16"]["stage1_ms"], 4), round(res["bf16"]["stage2_ms"], 4)],
                          "tile_cfg": res["bf16"]["tile_cfg"],
                          "frac_of_bf16_peak": round(res["bf16"]["tflops"] / PEAK_BF16, 4),
                          "rel_err_vs_fp64": rel16,
                          "speedup_vs_fp32": round(res["fp32"]["ms_per_step"] / res["bf16"]["ms_per_step"], 3)},
        "sweep": sweep,
        "roofline": {"bound": "mfma", "kernel": f"grouped_gemm {dom} (fp32): {dom_kernel}",
                     "achieved": round(achieved, 3),
                     "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32, 4),
                     "traffic": traffic,
                     "fp64_path": {"achieved": round(res["fp64"]["tflops"], 3), "peak": PEAK_FP64,
                                   "frac": round(res["fp64"]["tflops"] / PEAK_FP64, 4)}},
        "real_centre_bond": real,
        "sigma_variant": sig24,
        **extras,
        "cpu_baseline": cpu_summary(cpu, flops, real, sweep),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
