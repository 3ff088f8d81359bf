"""Every launch plan of the H_eff grouped GEMMs, forced one at a time
(sadmrg_heff_launch_plan): tile configurations 64x64, 128x128, 128x64 and
32x32, both split variants and both launch kinds, on a real SU(2) centre
bond whose stage-2 reductions split and on a complex U(1) one, against the
reference library's hamiltonian_vector_multiplication_idmrg
(dmrg_inst.cpp:283-371): FP64 to 1e-12, FP32 operands to 2e-5.  Complex
tables run tile configurations 0 and 3 on the complex kernel (other plans
are refused) and every plan as planar real tables."""
import numpy as np
import pytest

import sadmrg_testlib as tl

pytestmark = pytest.mark.gpu

N_PLANS = 16  # 4 tile configurations x 2 split variants x 2 launch kinds


@pytest.fixture
def plan_reset(sad):
    prev, prev_planar = sad.heff_launch_plan(), sad.heff_complex_planar()
    yield
    sad.heff_launch_plan(prev)
    sad.heff_complex_planar(prev_planar)


def _reference(wl, cgc2_kind):
    import refbind as rb
    if not rb.available():
        pytest.skip("oracle/_ref not built")
    d, il, ir = tl.reference_heff_setup(rb, wl, 2)
    return tl.from_ref(d.heff(il, ir, tl.to_ref(wl["psi"], cgc2_kind)))


def _check(sad, plan, psi, exp, plan_id):
    got = tl.from_dev(plan.apply(psi, sad.FP64))
    st = plan.stats()
    assert st["tile_cfg1"] == plan_id and st["tile_cfg2"] == plan_id, (plan_id, st)
    rel, _ = tl.compare(got, exp)
    assert rel <= 1e-12, (plan_id, rel)
    got32 = tl.from_dev(plan.apply(psi, sad.FP32))
    rel32, _ = tl.compare(got32, exp)
    assert rel32 <= 2e-5, (plan_id, rel32)


def test_every_launch_plan_real(ctx, sad, cgc, plan_reset):
    wl = tl.build_workload(D=256, n_op=3, seed=3)
    exp = _reference(wl, 1)
    psi = tl.to_dev(ctx, wl["psi"], 1)
    hl, hr = tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"])
    o1 = [tl.to_dev(ctx, t) for t in wl["o1"]]
    o2 = [tl.to_dev(ctx, t) for t in wl["o2"]]
    plan = sad.Heff(ctx, cgc, hl, hr, o1, o2, psi)
    for p in range(N_PLANS):
        sad.heff_launch_plan(p)
        _check(sad, plan, psi, exp, p)
    assert plan.stats()["split_tiles"] > 0


def test_every_launch_plan_complex(ctx, sad, plan_reset):
    wl = tl.build_workload_u1(m=128, length=32, seed=5, cplx=1)
    exp = _reference(wl, 0)
    psi = tl.to_dev(ctx, wl["psi"])
    plan = sad.Heff(ctx, None, tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"]),
                    [tl.to_dev(ctx, t) for t in wl["o1"]], [tl.to_dev(ctx, t) for t in wl["o2"]], psi)
    for planar in (False, True):
        sad.heff_complex_planar(planar)
        for p in range(N_PLANS):
            sad.heff_launch_plan(p)
            if not planar and p % 4 not in (0, 3):
                with pytest.raises(sad.SadmrgError):
                    plan.apply(psi, sad.FP64)
                continue
            _check(sad, plan, psi, exp, p)
