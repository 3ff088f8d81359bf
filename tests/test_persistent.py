"""Persistent launches of the grouped GEMM (sadmrg_gemm_persistent): a grid
of about the resident workgroup count takes output tiles from a device
ticket counter that the last workgroup resets.  Each tile is computed by
the same code in the same order as in the one-workgroup-per-tile launch,
so outputs are bitwise identical, launch after launch (the counter reset),
and from replayed Lanczos graphs."""
import os

import numpy as np
import pytest

import sadmrg_testlib as tl

pytestmark = pytest.mark.gpu


@pytest.fixture
def persistent_off(sad):
    prev = sad.persistent()
    yield
    sad.persistent(prev)


def _flat(t):
    b = tl.blocks(tl.from_dev(t))
    return np.concatenate([b[k].ravel() for k in sorted(b)])


@pytest.mark.parametrize("D,n_op", [(256, 4), (1024, 6)])
def test_persistent_heff_bitwise(ctx, sad, persistent_off, D, n_op):
    wl = tl.build_workload(D=D, n_op=n_op)
    cgc = sad.Cgc(1)
    psi = tl.to_dev(ctx, wl["psi"], 1)
    hl, hr = tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"])
    o1 = [tl.to_dev(ctx, t) for t in wl["o1"]]
    o2 = [tl.to_dev(ctx, t) for t in wl["o2"]]
    plan = sad.Heff(ctx, cgc, hl, hr, o1, o2, psi)
    for prec in (sad.FP64, sad.FP32, sad.BF16):
        sad.persistent(0)
        ref = _flat(plan.apply(psi, prec))
        sad.persistent(1)
        for rep in range(3):  # the ticket counter is reset by every launch
            got = _flat(plan.apply(psi, prec))
            assert np.array_equal(got, ref), (prec, rep, float(np.max(np.abs(got - ref))))


def test_persistent_sweep_identical(ctx, sad, persistent_off):
    """iDMRG + one sweep (environment updates, H_eff in graph-replayed
    Lanczos steps, truncation projectors) with persistent launches: every
    step energy equals the one-tile-per-workgroup run's."""
    tens, extra = tl.load_tensors(os.path.join(tl.GOLDEN, "dmrg_su2_obc12.npz"))
    sym, cplx, xx, yy, D, ns = (int(x) for x in extra["params"])
    runs = []
    for pers in (0, 1, 2):
        sad.persistent(pers)
        sad.lib().sadmrg_rng_seed(42)
        d = sad.DMRG(ctx, sym, cplx, xx, yy, D)
        d.set_couplings(extra["hmap"], extra["coup"])
        d.set_max_dcut(D)
        d.idmrg()
        d.sweep()
        runs.append(np.asarray([r["energy"] for r in d.records()]))
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])
