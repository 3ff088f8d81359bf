"""Generates the golden fixtures in tests/golden/ from the reference library.

Runs only in a container where oracle/_ref/libsadmrg_ref.so was built from
the reference sources (make -C oracle).  Every output is data: inputs and the
reference's outputs on them.  Usage:

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py dmrg_su2_ring10   # one case

Each DMRG case runs in a fresh process so the reference's global generator
(std::mt19937_64 seeded 42, dense_inst.cpp:23) starts from its initial state.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import sadmrg_testlib as tl  # noqa: E402
import refbind as rb  # noqa: E402

M = 20  # max_angm used by the reference driver (dmrg_inst.cpp:181)


# ---------------------------------------------------------------------------
def make_cgc():
    rb.init(physpn=1, max_angm=M, max_dcut=64, psize=1)
    L = rb.lib()
    P2, PS = 3, 9
    hv = np.zeros(M ** 3)
    L.ref_fac_hamilt_vec(rb._dptr(hv))
    out = {"hamilt_vec": hv.reshape(M, M, M)}
    for which, name in enumerate(["onsite_left", "onsite_rght", "pairup_left", "pairup_rght"]):
        t = np.zeros(M ** 4 * PS)
        L.ref_fac_table5(which, rb._dptr(t))
        nz = np.nonzero(t)[0]
        out[name + "_idx"] = nz.astype(np.int64)
        out[name + "_val"] = t[nz]
    for which, name in enumerate(["transformation_left", "transformation_rght"]):
        t = np.zeros(M ** 5 * P2)
        L.ref_fac_transformation(which, rb._dptr(t))
        nz = np.nonzero(t)[0]
        out[name + "_idx"] = nz.astype(np.int64)
        out[name + "_val"] = t[nz]
    cg = []
    for j1 in range(0, 7):
        for j2 in range(0, 7):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1, 2):
                for m1 in range(-j1, j1 + 1, 2):
                    for m2 in range(-j2, j2 + 1, 2):
                        m3 = m1 + m2
                        if abs(m3) > j3:
                            continue
                        cg.append([j1, m1, j2, m2, j3, m3, L.ref_cg(j1, m1, j2, m2, j3, m3)])
    out["cg"] = np.asarray(cg)
    out["max_angm"] = np.asarray(M)
    np.savez_compressed(os.path.join(HERE, "cgc_su2.npz"), **out)
    print("cgc_su2.npz", {k: v.shape for k, v in out.items()})


# ---------------------------------------------------------------------------
def _su2_bond(rng, d, js, dmax):
    return (d, list(js), [int(rng.integers(1, dmax + 1)) for _ in js])


def random_hvm_case(rng, sym, cplx):
    """psi (2-leg), op1/op2 (3-leg, spin-1 channel) with consistent bonds."""
    if sym == 1:
        js = sorted(rng.choice(np.arange(0, 9), size=int(rng.integers(2, 5)), replace=False).tolist())
        # psi sectors share parity so that spin-1 operators connect them
        par = int(rng.integers(0, 2))
        js = sorted({2 * (j // 2) + par for j in js})
        b0 = _su2_bond(rng, 1, js, 7)
        b1 = (-1, b0[1], [int(rng.integers(1, 8)) for _ in js])
        ch = (-1, [2], [1])
        psi = tl.random_tensor(rng, 1, cplx, 0, [b0, b1])
        op1 = tl.random_tensor(rng, 1, cplx, 0, [(-b0[0], b0[1], b0[2]), ch, b0], density=0.85)
        op2 = tl.random_tensor(rng, 1, cplx, 0, [(-b1[0], b1[1], b1[2]), ch, b1], density=0.85)
        hl = tl.random_tensor(rng, 1, cplx, 0, [(-b0[0], b0[1], b0[2]), b0])
        hr = tl.random_tensor(rng, 1, cplx, 0, [(-b1[0], b1[1], b1[2]), b1])
    else:
        qs = sorted(rng.choice(np.arange(-6, 7, 2), size=int(rng.integers(3, 6)), replace=False).tolist())
        b0 = (1, qs, [int(rng.integers(1, 7)) for _ in qs])
        b1 = (1, [-q for q in qs], [int(rng.integers(1, 7)) for _ in qs])
        ch1 = (-1, [-2, 0, 2], [1, 1, 1])
        ch2 = (1, [-2, 0, 2], [1, 1, 1])
        psi = tl.random_tensor(rng, 0, cplx, 0, [b0, b1])
        op1 = tl.random_tensor(rng, 0, cplx, 0, [(-b0[0], b0[1], b0[2]), ch1, b0], density=0.85)
        op2 = tl.random_tensor(rng, 0, cplx, 0, [(-b1[0], b1[1], b1[2]), ch2, b1], density=0.85)
        hl = tl.random_tensor(rng, 0, cplx, 0, [(-b0[0], b0[1], b0[2]), b0])
        hr = tl.random_tensor(rng, 0, cplx, 0, [(-b1[0], b1[1], b1[2]), b1])
    return psi, op1, op2, hl, hr


def make_hvm():
    rb.init(physpn=1, max_angm=M, max_dcut=64, psize=1)
    for sym, cplx, tag in [(1, 0, "su2_real"), (0, 0, "u1_real"), (0, 1, "u1_complex"), (1, 1, "su2_complex")]:
        rng = np.random.default_rng(1000 + 10 * sym + cplx)
        tens = {}
        for c in range(4):
            psi, op1, op2, hl, hr = random_hvm_case(rng, sym, cplx)
            cg2 = 1 if sym == 1 else 0
            rp, r1, r2 = tl.to_ref(psi, cg2), tl.to_ref(op1), tl.to_ref(op2)
            out = tl.from_ref(rb.hvm(rp, r1, r2))
            rh, rr = tl.to_ref(hl), tl.to_ref(hr)
            lpsi = tl.from_ref(rb.contract(rh, 0, rp, 0))
            psir = tl.from_ref(rb.contract(rp, 1, rr, 0))
            for k, v in dict(psi=psi, op1=op1, op2=op2, hl=hl, hr=hr, out=out, lpsi=lpsi, psir=psir).items():
                tens[f"c{c}_{k}"] = v
        tl.save_tensors(os.path.join(HERE, f"hvm_{tag}.npz"), tens, ncase=4)
        print(f"hvm_{tag}.npz", len(tens))


# ---------------------------------------------------------------------------
def snapshot(d, ns, exci=0):
    tens = {}
    for i in range(ns):
        for which in ("uu", "hh"):
            t = tl.from_ref(d.tensor(which, i))
            if int(t["header"][0]) > 0:
                tens[f"{which}_{i}"] = t
        for j in range(ns):
            t = tl.from_ref(d.tensor("opr", i, j))
            if int(t["header"][0]) > 0:
                tens[f"opr_{i}_{j}"] = t
        for j in range(exci):  # stored states: site tensors and overlap environments
            for which in ("orth", "ovlp"):
                t = tl.from_ref(d.tensor(which, i, j))
                if int(t["header"][0]) > 0:
                    tens[f"{which}_{i}_{j}"] = t
    for j in range(exci):
        tens[f"overlapvec_0_{j}"] = tl.from_ref(d.tensor("overlapvec", 0, j))
    return tens


def ref_sweep(d, ns, nfree, steps, capture=None, d_max=1 << 30, exci=0):
    """DMRG::sweep (dmrg_inst.cpp:874-943) step by step; returns per-step records."""
    recs = []

    def step(kind, i):
        if kind == "left" or kind == "return":
            il, ir = i - 2, i - 1
            d.call("wavefunc_transformation", i - 1, 1)
            d.call("right2", i, 1)
            vec = d.input_vector(il, ir)
            d.call("left1", il)
            d.call("right1", ir)
            if kind == "left":
                p1, p2 = (1.0, 0.0) if i > nfree + 2 else (0.0, 1.0)
            else:
                p1, p2 = 1.0, 0.0
        else:
            il, ir = i + 1, i + 2
            d.call("wavefunc_transformation", i + 1, 0)
            d.call("left2", i, 1)
            vec = d.input_vector(il, ir)
            d.call("left1", il)
            d.call("right1", ir)
            p1, p2 = (0.0, 1.0) if i < ns - nfree - 3 else (1.0, 0.0)
        cap = capture is not None and capture["when"] == (kind, i) and "done" not in capture
        if cap:
            capture["state"] = snapshot(d, ns, exci)
            capture["vec"] = tl.from_ref(vec)
            capture["heff"] = tl.from_ref(d.heff(il, ir, vec))
            capture["il"], capture["ir"], capture["p1"], capture["p2"] = il, ir, p1, p2
        e = d.lanczos(il, ir, vec)
        if cap:
            capture["lanczos_e"] = e
            capture["lanczos_vec"] = tl.from_ref(vec)
        w = d.svd(vec, il, ir, p1, p2)
        if cap:
            capture["wout"] = w[:int(np.count_nonzero(w))]
            capture["uu_l"] = tl.from_ref(d.tensor("uu", il))
            capture["uu_r"] = tl.from_ref(d.tensor("uu", ir))
            capture["done"] = True
        nw = int(np.count_nonzero(w))
        dcut = min(nw, d_max)
        recs.append((il, ir, e, float(np.sum(w[dcut:nw] ** 2))))

    for i in range(ns // 2, nfree + 1, -1):
        step("left", i)
    for i in range(nfree, ns - nfree - 2):
        step("right", i)
    for i in range(ns - nfree - 1, ns // 2, -1):
        step("return", i)
    return recs


def dmrg_case(name, sym, cplx, xx, yy, D, lattice, nsweep, capture_at):
    rb.init(physpn=1, max_angm=M, max_dcut=D, psize=1)
    d = rb.RefDMRG(sym, cplx, xx, yy, D, 1, 0)
    ns = d.ns
    if lattice is not None:
        hmap, coup = lattice
        d.set_couplings(hmap, coup)
    hmap, coup = d.couplings()
    d.idmrg()
    e_idmrg = d.energy()
    cap = {"when": tuple(capture_at)}
    sweeps = []
    for s in range(nsweep):
        recs = ref_sweep(d, ns, 2, None, cap if s == 1 else None, d_max=D)
        sweeps.append(recs)
    tens = {}
    for k, v in cap["state"].items():
        tens["S_" + k] = v
    for k in ("vec", "heff", "lanczos_vec", "uu_l", "uu_r"):
        tens[k] = cap[k]
    energies = np.asarray([[r[2] for r in recs] for recs in sweeps])
    trunc = np.asarray([[r[3] for r in recs] for recs in sweeps])
    steps = np.asarray([[r[0], r[1]] for r in sweeps[0]], np.int32)
    tl.save_tensors(os.path.join(HERE, f"{name}.npz"), tens, hmap=hmap, coup=coup,
                    params=np.asarray([sym, cplx, xx, yy, D, ns], np.int32), e_idmrg=e_idmrg,
                    energies=energies, trunc=trunc, steps=steps, cap_step=np.asarray([cap["il"], cap["ir"]], np.int32),
                    cap_p=np.asarray([cap["p1"], cap["p2"]]), lanczos_e=cap["lanczos_e"], wout=cap["wout"])
    print(name, "E_idmrg", e_idmrg, "sweep energies", energies[:, -1])


CASES = {
    # reference-native lattice: periodic snake (xx=1 gives a ring)
    "dmrg_su2_ring10": dict(sym=1, cplx=0, xx=1, yy=10, D=16, lattice=None, nsweep=2, capture_at=("right", 3)),
    "dmrg_u1_ring10": dict(sym=0, cplx=0, xx=1, yy=10, D=24, lattice=None, nsweep=2, capture_at=("right", 3)),
    # complex scalars (DMRG<complex<double>, U1Symmetry>), same lattice
    "dmrg_u1c_ring10": dict(sym=0, cplx=1, xx=1, yy=10, D=24, lattice=None, nsweep=2, capture_at=("right", 3)),
    # open chain and J1-J2 cylinder through the coupling map
    "dmrg_su2_obc12": dict(sym=1, cplx=0, xx=1, yy=12, D=24, lattice="obc", nsweep=2, capture_at=("left", 5)),
    "dmrg_su2_j1j2_4x3": dict(sym=1, cplx=0, xx=3, yy=4, D=24, lattice="j1j2", nsweep=2, capture_at=("right", 4)),
    # U(1) open chain L=32, m=64 kept states, four sweeps
    "dmrg_u1_obc32": dict(sym=0, cplx=0, xx=1, yy=32, D=64, lattice="obc", nsweep=4, capture_at=("right", 12)),
}


# ---------------------------------------------------------------------------
# Checkpoint files written by the reference's own sweep() (dmrg_inst.cpp:940-942:
# save_mps2, save_ww, save_enr into the working directory), and the resume the
# reference's driver performs from them (main/main_dmrg_su2_real.cpp:73-80:
# read_mps(read, read2), prepare_sweep, sweep).
CKPT_CASES = {
    "ckpt_su2_ring10": dict(xx=1, yy=10, D=16, lattice=None),
    "ckpt_su2_obc12": dict(xx=1, yy=12, D=24, lattice="obc"),
}


def ckpt_case(name, xx, yy, D, lattice):
    import shutil
    import tempfile
    rb.init(physpn=1, max_angm=M, max_dcut=D, psize=1)
    ns = xx * yy
    out_dir = os.path.join(HERE, name)
    if os.path.isdir(out_dir):
        shutil.rmtree(out_dir)
    os.makedirs(out_dir)
    cwd = os.getcwd()
    try:
        os.chdir(out_dir)
        d = rb.RefDMRG(1, 0, xx, yy, D, 1, 0)
        if lattice is not None:
            d.set_couplings(*lattice)
        hmap, coup = d.couplings()
        d.idmrg()
        d.sweep()  # writes uu-*, ww-*, enr-* into out_dir
        e_saved = d.energy()
        tens = {f"uu_{i}": tl.from_ref(d.tensor("uu", i)) for i in range(ns)}
        ww = d.ww(2 * D)
        with tempfile.TemporaryDirectory() as tmp:
            os.chdir(tmp)  # the resumed sweep writes its own files here
            r = rb.RefDMRG(1, 0, xx, yy, D, 1, 0)
            if lattice is not None:
                r.set_couplings(*lattice)
            ok = r.read(out_dir, D, 0)
            read_uu = {f"read_uu_{i}": tl.from_ref(r.tensor("uu", i)) for i in range(ns)}
            e_read = r.energy()
            r.call("prepare_sweep")
            r.sweep()
            e_resume = r.energy()
            os.chdir(out_dir)
    finally:
        os.chdir(cwd)
    tens.update(read_uu)
    tl.save_tensors(os.path.join(HERE, f"{name}.npz"), tens, hmap=hmap, coup=coup,
                    params=np.asarray([1, 0, xx, yy, D, ns], np.int32), e_saved=e_saved, e_read=e_read,
                    e_resume=e_resume, read_pass=int(ok), ww=ww)
    print(name, "E saved", e_saved, "read pass", ok, "E read", e_read, "E after resumed sweep", e_resume)


# Excited state exci = 1 (dmrg_inst.cpp:283-583) as the reference's driver
# runs it (main/main_dmrg_su2_real.cpp:72-96): read_mps(D, D) finds no files
# of state 1, loads state 0 (orth[0], gs_enr[0]) from the checkpoint fixture
# of the same lattice, then do_idmrg and sweeps.  Runs in a fresh process, so
# the generator starts from its initial state; the mid-sweep capture of the
# second sweep holds the stored state's tensors, its overlap environments and
# the overlap vector next to the usual state.
EXCI_CASES = {
    "exci_su2_ring10": dict(ckpt="ckpt_su2_ring10", xx=1, yy=10, D=16, lattice=None, nsweep=2,
                            capture_at=("right", 3)),
    "exci_su2_obc12": dict(ckpt="ckpt_su2_obc12", xx=1, yy=12, D=24, lattice="obc", nsweep=2,
                           capture_at=("left", 5)),
}


def exci_case(name, ckpt, xx, yy, D, lattice, nsweep, capture_at):
    import tempfile
    rb.init(physpn=1, max_angm=M, max_dcut=D, psize=1)
    ns = xx * yy
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)  # nothing is written here by the step-by-step sweeps; kept as a guard
        try:
            d = rb.RefDMRG(1, 0, xx, yy, D, 1, 0, exci=1)
            if lattice is not None:
                d.set_couplings(*lattice)
            hmap, coup = d.couplings()
            ok = d.read(os.path.join(HERE, ckpt), D, D)
            e_lower = d.state_energy(0)
            d.idmrg()
            e_idmrg = d.energy()
            cap = {"when": tuple(capture_at)}
            sweeps = []
            for s_ in range(nsweep):
                sweeps.append(ref_sweep(d, ns, 2, None, cap if s_ == 1 else None, d_max=D, exci=1))
        finally:
            os.chdir(cwd)
    tens = {"S_" + k: v for k, v in cap["state"].items()}
    for k in ("vec", "heff", "lanczos_vec", "uu_l", "uu_r"):
        tens[k] = cap[k]
    energies = np.asarray([[r[2] for r in recs] for recs in sweeps])
    steps = np.asarray([[r[0], r[1]] for r in sweeps[0]], np.int32)
    tl.save_tensors(os.path.join(HERE, f"{name}.npz"), tens, hmap=hmap, coup=coup,
                    params=np.asarray([1, 0, xx, yy, D, ns], np.int32), e_idmrg=e_idmrg, e_lower=e_lower,
                    read_pass=int(ok), energies=energies, steps=steps,
                    cap_step=np.asarray([cap["il"], cap["ir"]], np.int32),
                    cap_p=np.asarray([cap["p1"], cap["p2"]]), lanczos_e=cap["lanczos_e"])
    print(name, "read pass", ok, "E0", e_lower, "E_idmrg", e_idmrg, "sweep energies", energies[:, -1])


def run_exci(name):
    c = dict(EXCI_CASES[name])
    lat = c.pop("lattice")
    exci_case(name, lattice=lattice_of(lat, c["xx"], c["yy"]), **c)


def lattice_of(lat, xx, yy):
    ns = xx * yy
    if lat == "obc":
        return tl.open_chain(ns), np.asarray([0.0, 1.0] + [0.0] * (ns - 1))
    if lat == "j1j2":
        return tl.j1j2_cylinder(yy, xx), np.asarray([0.0, 1.0, 0.5] + [0.0] * (ns - 2))
    return None


def run_ckpt(name):
    c = dict(CKPT_CASES[name])
    lat = c.pop("lattice")
    ckpt_case(name, lattice=lattice_of(lat, c["xx"], c["yy"]), **c)


def run_case(name):
    c = dict(CASES[name])
    lat = c.pop("lattice")
    ns = c["xx"] * c["yy"]
    if lat == "obc":
        h = tl.open_chain(ns)
        lattice = (h, np.asarray([0.0, 1.0] + [0.0] * (ns - 1)))
    elif lat == "j1j2":
        h = tl.j1j2_cylinder(c["yy"], c["xx"])
        lattice = (h, np.asarray([0.0, 1.0, 0.5] + [0.0] * (ns - 2)))
    else:
        lattice = None
    dmrg_case(name, lattice=lattice, **c)


def main(argv):
    if len(argv) > 1 and argv[1] in CASES:
        run_case(argv[1])
        return
    if len(argv) > 1 and argv[1] in CKPT_CASES:
        run_ckpt(argv[1])
        return
    if len(argv) > 1 and argv[1] in EXCI_CASES:
        run_exci(argv[1])
        return
    if len(argv) > 1 and argv[1] == "cgc":
        make_cgc()
        return
    if len(argv) > 1 and argv[1] == "hvm":
        make_hvm()
        return
    for sub in ["cgc", "hvm"] + list(CASES) + list(CKPT_CASES) + list(EXCI_CASES):
        subprocess.run([sys.executable, __file__, sub], check=True)


if __name__ == "__main__":
    main(sys.argv)
