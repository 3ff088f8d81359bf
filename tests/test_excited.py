"""Excited states through the orthogonality penalty (dmrg_inst.cpp:283-583).

A driver built with max_exci = 1 keeps stored state 0 (orth[0], its energy
gs_enr[0]) and its overlap environments with the current MPS (ovlp[0]); on
two adjacent sites H_eff carries -E_0 |phi_0><phi_0|, phi_0 = state 0
projected on the two-site basis (overlapvec[0]).  Fixtures
tests/golden/exci_su2_*.npz hold the reference's run as its driver performs
it (main/main_dmrg_su2_real.cpp:72-96): read_mps(D, D) loads state 0 from
the checkpoint fixture of the same lattice, do_idmrg, two sweeps; the second
sweep's capture holds the state, the stored state's tensors and overlap
environments, the overlap vector, H_eff psi and the Lanczos energy.

The mathematics pins the result independently: the penalised spectrum's
lowest value is the first singlet excitation, from exact diagonalisation in
the total-spin-0 sector.
"""
import os

import numpy as np
import pytest

import sadmrg_testlib as tl

HERE = os.path.dirname(os.path.abspath(__file__))
EXCI = {"exci_su2_ring10": "ckpt_su2_ring10", "exci_su2_obc12": "ckpt_su2_obc12"}


def _case(name):
    tens, extra = tl.load_tensors(os.path.join(HERE, "golden", f"{name}.npz"))
    return tens, extra, os.path.join(HERE, "golden", EXCI[name])


def _singlets(extra, n=2):
    ns = int(extra["params"][5])
    return tl.heisenberg_ed_singlets(np.asarray(extra["hmap"]).reshape(ns, ns), extra["coup"], n)


# ---- CPU: the fixtures and the exact spectrum ------------------------------------------
def test_singlet_spectrum_brute_force():
    """heisenberg_ed_singlets against dense H and S^2 of a 6-site ring built
    from Kronecker products: the singlet energies are the eigenvalues of H
    restricted to the S^2 = 0 eigenspace."""
    ns = 6
    sx = np.array([[0, 0.5], [0.5, 0]])
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    sz = np.array([[0.5, 0], [0, -0.5]])

    def site(op, i):
        m = np.array([[1.0]])
        for k in range(ns):
            m = np.kron(m, op if k == i else np.eye(2))
        return m

    S = [[site(o, i) for o in (sx, sy, sz)] for i in range(ns)]
    dot = lambda i, j: sum(S[i][a] @ S[j][a] for a in range(3)).real  # noqa: E731
    hmap = np.asarray(tl.reference_lattice(1, ns))
    H = sum(dot(i, j) for i in range(ns) for j in range(i + 1, ns) if hmap[i, j] > 0)
    tot = [sum(S[i][a] for i in range(ns)) for a in range(3)]
    S2 = sum(t @ t for t in tot).real
    w, v = np.linalg.eigh(S2)
    P = v[:, np.abs(w) < 1e-9]
    sing = np.linalg.eigvalsh(P.T @ H @ P)
    got = tl.heisenberg_ed_singlets(hmap, np.array([0.0, 1.0]), 3)
    assert np.allclose(got, sing[:3], atol=1e-10), (got, sing[:3])


@pytest.mark.parametrize("name", list(EXCI))
def test_reference_excited_state_is_first_singlet(name):
    """The reference's exci = 1 run converges to the first singlet
    excitation (exact diagonalisation); its stored state's energy is the
    checkpoint's ground energy."""
    tens, extra, ckpt = _case(name)
    e0, e1 = _singlets(extra)
    assert int(extra["read_pass"]) == 0  # no files of state 1: do_idmrg follows
    assert abs(float(extra["e_lower"]) - e0) <= 1e-8 * abs(e0)
    last = float(extra["energies"][-1][-1])
    assert abs(last - e1) <= 1e-7 * abs(e1), (last, e1)


# ---- GPU ---------------------------------------------------------------------------
def _driver(sad, ctx, extra, exci=1):
    _, _, xx, yy, D, ns = (int(x) for x in extra["params"])
    d = sad.DMRG(ctx, 1, 0, xx, yy, D, exci=exci)
    d.set_couplings(extra["hmap"], extra["coup"])
    d.set_max_dcut(D)
    return d, D, ns


def _from_capture(sad, ctx, tens, extra, ckpt):
    """Driver holding the reference's captured mid-sweep state (state 0's
    energy from the checkpoint's enr file)."""
    d, D, ns = _driver(sad, ctx, extra)
    assert not d.read_mps(ckpt, D, D)
    for k, v in tens.items():
        if not k.startswith("S_"):
            continue
        parts = k[2:].split("_")
        if parts[0] in ("opr", "orth", "ovlp", "overlapvec"):
            d.set_tensor(parts[0], int(parts[1]), int(parts[2]), tl.to_dev(ctx, v))
        else:
            d.set_tensor(parts[0], int(parts[1]), 0, tl.to_dev(ctx, v))
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(EXCI))
def test_excited_state_run_matches_reference(ctx, sad, name):
    """read_mps(D, D) + do_idmrg + two sweeps with max_exci = 1: state 0 read
    bit for bit, every step energy of the converged second sweep equal to the
    reference's (1e-9), the result equal to the first singlet excitation."""
    tens, extra, ckpt = _case(name)
    ckt, _ = tl.load_tensors(os.path.join(HERE, "golden", f"{EXCI[name]}.npz"))
    import sadmrg_oracle as so
    e0, e1 = _singlets(extra)
    sad.lib().sadmrg_rng_seed(42)
    d, D, ns = _driver(sad, ctx, extra)
    assert not d.read_mps(ckpt, D, D)
    assert d.lower_energy(0) == float(extra["e_lower"])
    for i in range(ns):
        got = so.from_packed(tl.from_dev(d.tensor("orth", i, 0)))
        ref = so.from_packed(ckt[f"read_uu_{i}"])
        for b in set(got.blocks) | set(ref.blocks):
            x, y = got.blocks.get(b), ref.blocks.get(b)
            assert x is not None and y is not None and np.array_equal(x, y), (i, b)
    d.idmrg()
    ref0 = float(extra["e_idmrg"])
    assert abs(d.energy() - ref0) <= 1e-6 * abs(ref0), (d.energy(), ref0)
    d.records()
    energies = extra["energies"]
    for s in range(energies.shape[0]):
        d.sweep()
        got = np.asarray([r["energy"] for r in d.records()])
        assert got.shape == energies[s].shape
        # lowest eigenvalue of the penalised operator: the first singlet
        # excitation, up to the stored state's own error
        assert got.min() >= e1 - 1e-8 * abs(e1), (s, got.min(), e1)
    assert np.max(np.abs(got - energies[-1])) <= 1e-9 * np.max(np.abs(energies[-1])), (got, energies[-1])
    assert abs(d.energy() - e1) <= 1e-7 * abs(e1), (d.energy(), e1)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(EXCI))
def test_overlap_vector_and_penalised_heff_on_reference_state(ctx, sad, name):
    """prepare_input_vector builds overlapvec[0] equal to the reference's
    (dmrg_inst.cpp:429-436); H_eff psi including -E_0 phi <phi|psi>
    (dmrg_inst.cpp:352-363) equals the reference's in FP64 (1e-12) and
    within 1e-5 on the FP32-operand path; the Lanczos energy equals the
    reference's (1e-10)."""
    tens, extra, ckpt = _case(name)
    d = _from_capture(sad, ctx, tens, extra, ckpt)
    il, ir = (int(x) for x in extra["cap_step"])
    vec = d.input_vector(il, ir)
    rel, _ = tl.compare(tl.from_dev(vec), tens["vec"])
    assert rel <= 1e-12, rel
    ov = tl.from_dev(d.tensor("overlapvec", 0, 0))
    exp = tens["S_overlapvec_0_0"]
    assert list(ov["header"]) == list(exp["header"])
    rel, _ = tl.compare(ov, exp)
    assert rel <= 1e-12, rel
    v = tl.to_dev(ctx, tens["vec"], 1)
    rel, _ = tl.compare(tl.from_dev(d.heff(il, ir, v)), tens["heff"])
    assert rel <= 1e-12, rel
    rel32, _ = tl.compare(tl.from_dev(d.heff(il, ir, v, sad.FP32)), tens["heff"])
    assert rel32 <= 1e-5, rel32
    e = d.lanczos(il, ir, tl.to_dev(ctx, tens["vec"], 1))
    e_ref = float(extra["lanczos_e"])
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref), (e, e_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(EXCI))
def test_overlap_environment_updates_replay_reference(ctx, sad, name):
    """prepare_site_operator_from_{left,right}(i, 1) on the captured state
    recompute ovlp[0][i] and the two 3-leg pieces next to the pair
    (dmrg_inst.cpp:493-500, 567-574) equal to the reference's."""
    tens, extra, ckpt = _case(name)
    il, ir = (int(x) for x in extra["cap_step"])
    ns = int(extra["params"][5])
    assert il - 1 >= 1 and ir + 1 <= ns - 2
    checked = 0
    for call, site, sites in (("left2", il - 1, (il - 1, il, ir)), ("right2", ir + 1, (ir + 1, il, ir))):
        d = _from_capture(sad, ctx, tens, extra, ckpt)
        d.call(call, site, 1)
        for s in sites:
            rel, _ = tl.compare(tl.from_dev(d.tensor("ovlp", s, 0)), tens[f"S_ovlp_{s}_0"])
            assert rel <= 1e-12, (call, site, s, rel)
            checked += 1
    assert checked == 6


@pytest.mark.gpu
def test_lower_state_from_driver(ctx, sad):
    """set_lower_state (state 0 from a converged driver in memory, no files)
    gives the same excited state as the file path: the first singlet
    excitation of the 10-site ring; a clone of the excited driver continues
    to the same energy; the coupling gradient refuses penalised drivers."""
    tens, extra, _ = _case("exci_su2_ring10")
    e0, e1 = _singlets(extra)
    sad.lib().sadmrg_rng_seed(42)
    g, D, ns = _driver(sad, ctx, extra, exci=0)
    g.idmrg()
    g.sweep()
    g.sweep()
    assert abs(g.energy() - e0) <= 1e-9 * abs(e0)
    x, _, _ = _driver(sad, ctx, extra)
    x.set_lower_state(0, g)
    assert x.lower_energy(0) == g.energy()
    x.idmrg()
    x.sweep()
    y = x.clone()
    x.sweep()
    y.sweep()
    assert abs(x.energy() - e1) <= 1e-7 * abs(e1), (x.energy(), e1)
    assert abs(y.energy() - x.energy()) <= 1e-12 * abs(e1)
    with pytest.raises(sad.SadmrgError, match="ground-state"):
        x.coupling_gradient()


@pytest.mark.gpu
def test_excited_checkpoint_files(ctx, sad, tmp_path):
    """A max_exci = 1 driver names its files with exci = 1 (save_mps2(exci),
    save_enr; dmrg_inst.cpp:940-1003); a second driver resumes from them
    (read_mps finds its own files and the stored state's) to the same energy;
    a missing stored state is an error."""
    tens, extra, ckpt = _case("exci_su2_ring10")
    e0, e1 = _singlets(extra)
    sad.lib().sadmrg_rng_seed(42)
    d, D, ns = _driver(sad, ctx, extra)
    assert not d.read_mps(ckpt, D, D)
    d.idmrg()
    import shutil
    for f in os.listdir(ckpt):
        shutil.copy(os.path.join(ckpt, f), tmp_path / f)
    d.set_checkpoint_dir(str(tmp_path))
    d.sweep()
    d.set_checkpoint_dir(None)
    assert all((tmp_path / f"uu-{ns}-{D}-0-1-{i}.bin").exists() for i in range(ns))
    with open(tmp_path / f"enr-{ns}-{D}-0-1.dat") as f:
        assert abs(float(f.read().split()[-1]) - d.energy()) <= 1e-15 * abs(d.energy())
    r, _, _ = _driver(sad, ctx, extra)
    assert r.read_mps(str(tmp_path), D, D)
    assert r.energy() == float("%.16g" % d.energy())
    r.prepare_sweep()
    d.prepare_sweep()
    r.sweep()
    d.sweep()
    assert abs(r.energy() - d.energy()) <= 1e-12 * abs(e1)
    empty = tmp_path / "empty"
    empty.mkdir()
    q, _, _ = _driver(sad, ctx, extra)
    with pytest.raises(sad.SadmrgError, match="stored state"):
        q.read_mps(str(empty), D, D)
