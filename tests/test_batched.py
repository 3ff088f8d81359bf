"""Instances batched into one launch (HeffPlan::merge): the block tables of
independent H_eff plans -- chains of different lengths, here with U(1)
complex operands as in the XXZ + Dzyaloshinskii-Moriya batch -- merged so
that each stage of every instance runs in a single grouped launch.  No
reference counterpart: the reference applies one H_eff at a time
(dmrg_inst.cpp:283-371), which is what each instance is checked against.
"""
import numpy as np
import pytest

import sadmrg_testlib as tl


def _instances(n, m, seed=7, cplx=1):
    rng = np.random.default_rng(seed)
    lengths = [int(x) for x in rng.choice([32, 48, 64, 96], size=n)]
    return [tl.build_workload_u1(m=m, length=L, seed=seed + 1 + i, cplx=cplx) for i, L in enumerate(lengths)]


def test_u1_workload_structure():
    """The synthetic U(1) centre bond: psi holds exactly the S^z = 0 blocks
    (equal fused labels on both legs), its length is the sum of squared
    fused dimensions, the block Hamiltonians are Hermitian, and wider
    chains give wider 2S^z histograms."""
    wl = tl.build_workload_u1(m=128, length=48, seed=3)
    _, _, bonds = tl.parse_header(wl["psi"]["header"])
    fq, fd = wl["info"]["two_sz_fused"], wl["info"]["dims_fused"]
    present = [tl.block_labels(bonds, b)[0] for b in range(tl.nten(bonds)) if wl["psi"]["present"][b]]
    assert sorted(p[0] for p in present) == sorted(fq) and all(p[0] == p[1] for p in present)
    assert wl["psi"]["data"].size == 2 * sum(d * d for d in fd)
    assert sum(wl["info"]["dims_block"]) in range(120, 137)
    blocks = tl.blocks(wl["hl"])
    for v in blocks.values():
        assert np.allclose(v, v.conj().T)
    s32 = tl.build_workload_u1(m=128, length=32, seed=3)["info"]["sigma"]
    s96 = tl.build_workload_u1(m=128, length=96, seed=3)["info"]["sigma"]
    assert s96 > s32


def _plans(ctx, sad, wls):
    plans, psis = [], []
    for wl in wls:
        psi = tl.to_dev(ctx, wl["psi"])
        o1 = [tl.to_dev(ctx, t) for t in wl["o1"]]
        o2 = [tl.to_dev(ctx, t) for t in wl["o2"]]
        plans.append(sad.Heff(ctx, None, tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"]), o1, o2, psi))
        psis.append(psi)
    return plans, psis


def _concat(ctx, sad, wls):
    """All instances' psi back to back as one device tensor (a flat U(1)
    block of the total length) and the offsets of each instance."""
    data = np.concatenate([wl["psi"]["data"] for wl in wls])
    n = data.size // 2
    t = tl.to_dev(ctx, dict(sym=0, cplx=1, header=tl.make_header(0, [(1, [0], [n]), (-1, [0], [1])]),
                            present=np.ones(1, np.uint8), data=data))
    return t, n


def _flat(t):
    return tl.from_dev(t)["data"]


@pytest.mark.gpu
def test_merged_plan_equals_individual_plans(ctx, sad):
    """One merged launch per stage gives, instance by instance, the output of
    that instance's own plan: bit for bit on the complex kernel (every output
    tile is reduced by one workgroup in segment order either way), to 1e-13
    on the planar real tables (the merged table may split long reductions
    differently); the complex64 operand path stays within 1e-5 of FP64 and
    differs from it."""
    wls = _instances(6, 64)
    plans, psis = _plans(ctx, sad, wls)
    merged = sad.Heff.merge(plans)
    assert merged.stats()["n"] == sum(wl["psi"]["data"].size // 2 for wl in wls)
    vin, n = _concat(ctx, sad, wls)
    out = vin.copy()
    merged.apply_device(vin.device_ptr(), out.device_ptr(), sad.FP64)
    got = _flat(out)
    off = 0
    planar = sad.heff_complex_planar()
    for plan, psi, wl in zip(plans, psis, wls):
        ref = tl.from_dev(plan.apply(psi, sad.FP64))["data"]
        k = ref.size
        if planar:
            assert np.max(np.abs(got[off:off + k] - ref)) <= 1e-13 * np.max(np.abs(ref))
        else:
            assert np.array_equal(got[off:off + k], ref)
        off += k
    out32 = vin.copy()
    merged.apply_device(vin.device_ptr(), out32.device_ptr(), sad.FP32)
    g32 = _flat(out32)
    rel = np.max(np.abs(g32 - got)) / np.max(np.abs(got))
    assert 1e-9 < rel <= 1e-5, rel
    with pytest.raises(sad.SadmrgError, match="merged"):
        merged.apply(psis[0])


@pytest.mark.gpu
def test_merged_plan_matches_reference_per_instance(ctx, sad):
    """Each instance's slice of the merged output equals the reference's
    hamiltonian_vector_multiplication_idmrg (DMRG<complex<double>, U1>) on
    the same tensors (1e-12)."""
    import refbind as rb
    if not rb.available():
        pytest.skip("oracle/_ref not built")
    wls = _instances(3, 48, seed=11)
    plans, _ = _plans(ctx, sad, wls)
    merged = sad.Heff.merge(plans)
    vin, n = _concat(ctx, sad, wls)
    out = vin.copy()
    merged.apply_device(vin.device_ptr(), out.device_ptr(), sad.FP64)
    got = _flat(out)
    off = 0
    for wl in wls:
        d, il, ir = tl.reference_heff_setup(rb, wl, 1)
        ref = tl.from_ref(d.heff(il, ir, tl.to_ref(wl["psi"])))
        mine = dict(wl["psi"])
        mine["data"] = got[off:off + wl["psi"]["data"].size]
        rel, _ = tl.compare(mine, ref)
        assert rel <= 1e-12, rel
        off += wl["psi"]["data"].size


@pytest.mark.gpu
def test_merged_real_plans(ctx, sad):
    """Real SU(2) centre bonds of different sizes merged: each slice equals
    its own plan's output to 1e-13 (the merged table may split long
    reductions differently)."""
    wls = [tl.build_workload(D=d, n_op=2, seed=s) for d, s in ((96, 1), (160, 2), (64, 3))]
    plans, psis = [], []
    for wl in wls:
        psi = tl.to_dev(ctx, wl["psi"], 1)
        o1 = [tl.to_dev(ctx, t) for t in wl["o1"]]
        o2 = [tl.to_dev(ctx, t) for t in wl["o2"]]
        plans.append(sad.Heff(ctx, sad.Cgc(1), tl.to_dev(ctx, wl["hl"]), tl.to_dev(ctx, wl["hr"]), o1, o2, psi))
        psis.append(psi)
    merged = sad.Heff.merge(plans)
    data = np.concatenate([wl["psi"]["data"] for wl in wls])
    vin = tl.to_dev(ctx, dict(sym=0, cplx=0, header=tl.make_header(0, [(1, [0], [data.size]), (-1, [0], [1])]),
                              present=np.ones(1, np.uint8), data=data))
    out = vin.copy()
    merged.apply_device(vin.device_ptr(), out.device_ptr(), sad.FP64)
    got = _flat(out)
    off = 0
    for plan, psi, wl in zip(plans, psis, wls):
        ref = tl.from_dev(plan.apply(psi, sad.FP64))["data"]
        assert np.max(np.abs(got[off:off + ref.size] - ref)) <= 1e-13 * np.max(np.abs(ref))
        off += ref.size


def _chain_instances(ctx, sad, lengths, m, seed=11):
    """XXZ + Dzyaloshinskii-Moriya chains (U(1) complex drivers), iDMRG and
    one sweep at m kept states each; the centre-bond H_eff plan of each and a
    seeded start vector (the sweep's input vector plus Gaussian noise)."""
    rng = np.random.default_rng(seed)
    plans, starts, psis = [], [], []
    for L in lengths:
        hmap = np.asarray(sad.open_chain(L), dtype=np.int32)
        coup = np.zeros(L + 1)
        coup[1] = 1.0
        d = sad.DMRG(ctx, sad.U1, 1, 1, L, m)
        d.set_couplings(hmap, coup)
        d.set_channel_weights(sad.xxz_dm_weights(L + 1, {1: (1.0, rng.uniform(0.5, 1.5), rng.uniform(0, 0.5))}))
        d.set_max_dcut(m)
        d.idmrg()
        d.sweep()
        il, ir = L // 2 - 1, L // 2
        psi = d.input_vector(il, ir)
        h = tl.from_dev(psi)
        h["data"] = h["data"] + 0.3 * np.abs(h["data"]).max() * rng.standard_normal(h["data"].size)
        starts.append(h)
        psis.append(psi)
        plans.append(d.heff_plan(il, ir, psi))
    return plans, starts, psis


def _dense(ctx, sad, plan, layout):
    """The plan's H_eff as a dense complex matrix (applied to unit vectors)."""
    h = tl.from_dev(layout)
    n = h["data"].size // 2
    cols = []
    for k in range(n):
        e = dict(h)
        e["data"] = np.zeros_like(h["data"])
        e["data"][2 * k] = 1.0
        o = tl.from_dev(plan.apply(tl.to_dev(ctx, e), sad.FP64))["data"]
        cols.append(o[0::2] + 1j * o[1::2])
    return np.stack(cols, axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_lanczos_matches_single_solves(ctx, sad, precision):
    """sadmrg_lanczos_batched on the merged plan of four chains of different
    lengths.  FP64: every instance follows its own single solve (the same
    lanczos1 schedule): energies to 1e-12, Ritz vectors to 1e-8 up to a
    phase.  Mixed (FP32 Krylov, FP64 checks, residual-correction restarts to
    |r| <= 1e-9 |E|): every energy equals the exact lowest eigenvalue of the
    instance's H_eff (dense, Hermitian) to 1e-11."""
    lengths = [8, 12, 10, 12]
    plans, starts, psis = _chain_instances(ctx, sad, lengths, m=16)
    merged = sad.Heff.merge(plans)
    ref_e, ref_v = [], []
    for p, s, psi in zip(plans, starts, psis):
        v = tl.to_dev(ctx, s)
        e, _ = p.lanczos(v, mlanc=40, precision=sad.FP64)
        ref_e.append(e)
        ref_v.append(tl.from_dev(v)["data"])
    data = np.concatenate([s["data"] for s in starts])
    n = data.size // 2
    vin = tl.to_dev(ctx, dict(sym=0, cplx=1, header=tl.make_header(0, [(1, [0], [n]), (-1, [0], [1])]),
                              present=np.ones(1, np.uint8), data=data))
    opts = dict(mlanc=40, precision=sad.FP64)
    if precision == "mixed":
        opts.update(precision=sad.FP32, refine=1, refine_tol=1e-9, refine_krylov=4, max_restarts=8)
    e, info = merged.lanczos_batched(vin.device_ptr(), len(plans), **opts)
    ref_e = np.asarray(ref_e)
    if precision == "fp64":
        assert np.all(np.abs(e - ref_e) <= 1e-12 * np.abs(ref_e)), (e - ref_e)
        got = tl.from_dev(vin)["data"]
        off = 0
        for i, r in enumerate(ref_v):
            g = got[off:off + r.size]
            off += r.size
            zg, zr = g[0::2] + 1j * g[1::2], r[0::2] + 1j * r[1::2]
            ph = np.vdot(zr, zg)
            ph /= abs(ph)
            assert np.max(np.abs(zg - ph * zr)) <= 1e-8, (i, float(np.max(np.abs(zg - ph * zr))))
    else:
        assert info["matvec32"] > 0
        assert np.all(info["residual64"] <= 1e-9 * np.abs(e) * (1 + 1e-6)), info["residual64"]
        exact = np.asarray([np.linalg.eigvalsh(_dense(ctx, sad, p, psi))[0] for p, psi in zip(plans, psis)])
        assert np.all(np.abs(e - exact) <= 1e-11 * np.abs(exact)), (e - exact)


@pytest.mark.gpu
def test_complex_planar_matches_complex_kernel(ctx, sad):
    """Complex plans lowered to real tables on planar operands (the default)
    against the complex kernel on interleaved data, single and merged plans,
    complex128 (1e-13) and complex64 operands (each within 1e-5 of complex128,
    and the two complex64 paths within 2e-6 of each other); the interleaved
    path stays bitwise reproducible across the switch."""
    wls = _instances(4, 64, seed=5)
    plans, psis = _plans(ctx, sad, wls)
    merged = sad.Heff.merge(plans)
    vin, n = _concat(ctx, sad, wls)
    prev = sad.heff_complex_planar()
    try:
        res = {}
        for planar in (False, True):
            sad.heff_complex_planar(planar)
            for prec in (sad.FP64, sad.FP32):
                out = vin.copy()
                merged.apply_device(vin.device_ptr(), out.device_ptr(), prec)
                single = np.concatenate([tl.from_dev(p.apply(a, prec))["data"] for p, a in zip(plans, psis)])
                res[(planar, prec)] = (_flat(out), single)
        ref = res[(False, sad.FP64)][1]
        scale = np.max(np.abs(ref))
        for planar in (False, True):
            for k in (0, 1):
                assert np.max(np.abs(res[(planar, sad.FP64)][k] - ref)) <= 1e-13 * scale
                assert np.max(np.abs(res[(planar, sad.FP32)][k] - ref)) <= 1e-5 * scale
        assert np.max(np.abs(res[(True, sad.FP32)][0] - res[(False, sad.FP32)][0])) <= 2e-6 * scale
        sad.heff_complex_planar(False)
        out = vin.copy()
        merged.apply_device(vin.device_ptr(), out.device_ptr(), sad.FP64)
        assert np.array_equal(_flat(out), res[(False, sad.FP64)][0])
    finally:
        sad.heff_complex_planar(prev)
