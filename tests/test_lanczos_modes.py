"""Lanczos step modes: host-driven steps, device steps (alpha, beta and the
step index on the device, the host synchronising only at convergence
checks) and device steps replayed from a HIP graph captured once per solve.

All three run the same arithmetic (dmrg_inst.cpp:1286-1348 restated in
lanczos.cpp: the same kernels, explicit fused multiply-adds, correctly
rounded division and square root), so energies, Ritz vectors and matvec
counts agree bit for bit.  Graph replays are interleaved with the solver's
own eager work (Ritz vectors, check matvecs, host synchronisation), and
every replayed step feeds the energy that is compared.
"""
import os

import numpy as np
import pytest

import sadmrg_testlib as tl

pytestmark = pytest.mark.gpu

MODES = {"host": (False, False), "device": (True, False), "graph": (True, True)}


@pytest.fixture
def restore_mode(sad):
    prev = sad.lanczos_mode()
    yield
    sad.lanczos_mode(prev[0], prev[1])


def _capture(name):
    return tl.load_tensors(os.path.join(tl.GOLDEN, f"{name}.npz"))


def _engine(ctx, sad, tens, extra):
    sym, cplx, xx, yy, D, ns = (int(x) for x in extra["params"])
    d = sad.DMRG(ctx, sym, cplx, xx, yy, D)
    d.set_couplings(extra["hmap"], extra["coup"])
    d.set_max_dcut(D)
    for k, v in tens.items():
        if not k.startswith("S_"):
            continue
        parts = k[2:].split("_")
        if parts[0] == "opr":
            d.set_tensor("opr", int(parts[1]), int(parts[2]), tl.to_dev(ctx, v))
        else:
            d.set_tensor(parts[0], int(parts[1]), 0, tl.to_dev(ctx, v))
    return d


def _flat(t):
    b = tl.blocks(tl.from_dev(t))
    return np.concatenate([b[k].ravel() for k in sorted(b)])


@pytest.mark.parametrize("name", ["dmrg_su2_ring10", "dmrg_u1_obc32", "dmrg_su2_j1j2_4x3"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lanczos_modes_agree_on_reference_state(ctx, sad, restore_mode, name, precision):
    """One Lanczos solve on the reference's captured mid-sweep state in each
    mode: identical energy and Ritz vector; the graph mode replays captured
    steps."""
    tens, extra = _capture(name)
    il, ir = (int(x) for x in extra["cap_step"])
    cg2 = 1 if tens["vec"]["sym"] == 1 else 0
    res = {}
    for mode, (dev, graph) in MODES.items():
        sad.lanczos_mode(dev, graph, reset=True)
        d = _engine(ctx, sad, tens, extra)
        if precision == "mixed":
            d.set_precision(sad.FP32, refine=True, refine_tol=1e-9)
        vec = tl.to_dev(ctx, tens["vec"], cg2)
        e = d.lanczos(il, ir, vec)
        counts = sad.lanczos_mode()
        res[mode] = (e, _flat(vec), counts[2])
    e_ref = float(extra["lanczos_e"])
    e0, v0, _ = res["host"]
    assert abs(e0 - e_ref) <= 1e-10 * abs(e_ref)
    for mode in ("device", "graph"):
        e, v, _ = res[mode]
        assert e == e0, (mode, e, e0)
        assert np.array_equal(v, v0), (mode, float(np.max(np.abs(v - v0))))
    assert res["graph"][2] > 0, "no Krylov step was replayed from a graph"
    assert res["device"][2] == 0 and res["host"][2] == 0


@pytest.mark.parametrize("name", ["dmrg_su2_obc12", "dmrg_u1_ring10"])
def test_full_run_step_energies_agree_across_modes(ctx, sad, restore_mode, name):
    """iDMRG + one sweep from the seeded start in each mode: every step
    energy and matvec count identical; the early iDMRG steps, whose
    Krylov spaces are exhausted after a few vectors, exercise the device-side
    breakdown stop."""
    tens, extra = _capture(name)
    sym, cplx, xx, yy, D, ns = (int(x) for x in extra["params"])
    runs = {}
    for mode, (dev, graph) in MODES.items():
        sad.lanczos_mode(dev, graph, reset=True)
        sad.lib().sadmrg_rng_seed(42)
        d = sad.DMRG(ctx, sym, cplx, xx, yy, D)
        d.set_couplings(extra["hmap"], extra["coup"])
        d.set_max_dcut(D)
        d.idmrg()
        d.sweep()
        recs = d.records()
        runs[mode] = (np.asarray([r["energy"] for r in recs]), np.asarray([r["matvec64"] for r in recs]),
                      sad.lanczos_mode()[2])
    e0, m0, _ = runs["host"]
    for mode in ("device", "graph"):
        e, m, _ = runs[mode]
        assert e.shape == e0.shape
        assert np.array_equal(e, e0), (mode, float(np.max(np.abs(e - e0))))
        assert np.array_equal(m, m0), (mode, m, m0)
    assert runs["graph"][2] > 0


@pytest.mark.parametrize("every", [0, 3])
def test_fp64_energy_policy(ctx, sad, restore_mode, every):
    """Mixed sweeps computing the FP64 energy at every `every`-th step and the
    last (every = 0: the last only).  The other steps make no FP64 matvec
    and record the FP32 Ritz value; the last step's FP64 energy matches the
    FP64 sweep's to 1e-9, the FP32 Ritz values to the FP32 operator error.
    Host-driven and graph-replayed steps agree bit for bit."""
    tens, extra = _capture("dmrg_su2_obc12")
    sym, cplx, xx, yy, D, ns = (int(x) for x in extra["params"])

    def run(prec, mode, e64every):
        sad.lanczos_mode(*MODES[mode], reset=True)
        sad.lib().sadmrg_rng_seed(42)
        d = sad.DMRG(ctx, sym, cplx, xx, yy, D)
        d.set_couplings(extra["hmap"], extra["coup"])
        d.set_max_dcut(D)
        if prec == "mixed":
            d.set_precision(sad.FP32, refine=True, refine_tol=1e-6)
            d.set_check_interval(2)
            d.set_fp64_energy_every(e64every)
        d.idmrg()
        d.sweep()
        d.sweep()
        return d.records()

    ref = run("fp64", "graph", 1)
    got = run("mixed", "graph", every)
    host = run("mixed", "host", every)
    assert len(got) == len(ref) == len(host)
    e_ref = np.asarray([r["energy"] for r in ref])
    e = np.asarray([r["energy"] for r in got])
    assert np.array_equal(e, np.asarray([r["energy"] for r in host]))
    # records: the iDMRG steps, then the two sweeps' steps
    nsweep = 2 * ns - 12
    n_idmrg = len(got) - 2 * nsweep
    flags = np.asarray([int(r["energy64"]) for r in got])
    assert np.all(flags[:n_idmrg] == 1)
    for s in range(2):
        f = flags[n_idmrg + s * nsweep:n_idmrg + (s + 1) * nsweep]
        want = np.zeros(nsweep, int)
        want[-1] = 1
        if every > 1:
            want[every - 1::every] = 1
        assert np.array_equal(f, want), (s, f.tolist())
    m64 = np.asarray([int(r["matvec64"]) for r in got])
    assert np.all(m64[flags == 0] == 0) and np.all(m64[flags == 1] >= 1)
    rel = np.abs(e - e_ref) / np.abs(e_ref)
    # the converged second sweep: FP64-energy steps to 1e-9, the others to
    # the FP32 operator error
    last = slice(n_idmrg + nsweep, None)
    f2 = flags[last] == 1
    assert np.all(rel[last][f2] <= 1e-9), rel[last][f2]
    assert np.all(rel[last][~f2] <= 1e-5), rel[last][~f2].max()
    r32 = np.asarray([r["ritz32"] for r in got])
    assert np.all(np.abs(r32[flags == 0] - e[flags == 0]) == 0)
