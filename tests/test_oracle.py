"""Pins the oracle (oracle/sadmrg_oracle.py) before it is trusted:
reference golden vectors, the reference library itself where it is built,
closed forms and exact diagonalisation.  CPU only."""
import os

import numpy as np
import pytest

import sadmrg_testlib as tl
import sadmrg_oracle as orc

HVM_TAGS = ["su2_real", "u1_real", "u1_complex", "su2_complex"]
DMRG_CASES = ["dmrg_su2_ring10", "dmrg_u1_ring10", "dmrg_u1c_ring10", "dmrg_su2_obc12", "dmrg_su2_j1j2_4x3",
              "dmrg_u1_obc32"]


# ---- closed forms -------------------------------------------------------------------
def test_cg_orthonormality_and_known_values():
    for j1 in range(0, 5):
        for j2 in range(0, 5):
            # sum_{m1 m2} C(j3 m3) C(j3' m3') = delta
            for j3 in range(abs(j1 - j2), j1 + j2 + 1, 2):
                for j3b in range(abs(j1 - j2), j1 + j2 + 1, 2):
                    for m3 in range(-j3, j3 + 1, 2):
                        s = sum(orc.cg(j1, m1, j2, m3 - m1, j3, m3) * orc.cg(j1, m1, j2, m3 - m1, j3b, m3)
                                for m1 in range(-j1, j1 + 1, 2) if abs(m3 - m1) <= j2)
                        if abs(m3) <= j3b:
                            assert abs(s - (1.0 if j3 == j3b else 0.0)) < 1e-13
    assert abs(orc.cg(1, 1, 1, -1, 2, 0) - np.sqrt(0.5)) < 1e-15
    assert abs(orc.cg(1, 1, 1, -1, 0, 0) - np.sqrt(0.5)) < 1e-15
    assert abs(orc.cg(1, -1, 1, 1, 0, 0) + np.sqrt(0.5)) < 1e-15
    assert abs(orc.cg(2, 2, 1, -1, 1, 1) - np.sqrt(2.0 / 3.0)) < 1e-15


def test_cg_against_reference_values():
    z = np.load(os.path.join(tl.GOLDEN, "cgc_su2.npz"))
    for row in z["cg"]:
        j1, m1, j2, m2, j3, m3 = (int(x) for x in row[:6])
        assert abs(orc.cg(j1, m1, j2, m2, j3, m3) - row[6]) <= 1e-14


def test_fac_tables_against_reference():
    z = np.load(os.path.join(tl.GOLDEN, "cgc_su2.npz"))
    M = int(z["max_angm"])
    hv = z["hamilt_vec"]
    for i in range(M):
        for k in range(M):
            assert abs(orc.fac_hamilt_vec(i, 2, k) - hv[i, 2, k]) <= 1e-12
            assert abs(orc.fac_hamilt_vec(i, 0, k) - hv[i, 0, k]) <= 1e-12
    for side, name in enumerate(["onsite_left", "onsite_rght"]):
        for idx, val in zip(z[name + "_idx"], z[name + "_val"]):
            j, m, l, i, p = (int(x) for x in np.unravel_index(idx, (M, M, M, M, 9)))
            assert abs(orc.fac_onsite(side, j, m, l, i, p) - val) <= 1e-12 * max(1, abs(val))
    for side, name in enumerate(["pairup_left", "pairup_rght"]):
        for idx, val in zip(z[name + "_idx"], z[name + "_val"]):
            j, i, k, m, p = (int(x) for x in np.unravel_index(idx, (M, M, M, M, 9)))
            assert abs(orc.fac_pairup(side, j, i, k, m, p) - val) <= 1e-12 * max(1, abs(val))
    for side, name in enumerate(["transformation_left", "transformation_rght"]):
        vals = z[name + "_val"]
        for idx, val in list(zip(z[name + "_idx"], vals))[::3]:
            j, m, l, i, k, p = (int(x) for x in np.unravel_index(idx, (M, M, M, M, M, 3)))
            assert abs(orc.fac_transformation(side, j, m, l, i, k, p) - val) <= 1e-12 * max(1, abs(val))


def test_fac_hamilt_vec_support_and_singlet_energy():
    """Two spin-1/2 sites in a singlet: the bond term sigma0 x sigma1 with the
    reduced elements +-sqrt(3)/2 (dmrg_inst.cpp:236-239) and weight
    fac_hamilt_vec[1][2][1] must give |<S1.S2>| = 3/4; the weight vanishes off
    the spin-1 triangle."""
    f = orc.fac_hamilt_vec(1, 2, 1)
    assert abs(abs((np.sqrt(3) / 2) * (np.sqrt(3) / 2) * f) - 0.75) < 1e-14
    for i in range(10):
        for k in range(10):
            if k < abs(i - 2) or k > i + 2 or (i + k) % 2:
                assert orc.fac_hamilt_vec(i, 2, k) == 0.0


# ---- block operations vs reference golden vectors ----------------------------------
@pytest.mark.parametrize("tag", HVM_TAGS)
def test_oracle_hvm_and_contract(tag):
    t, e = tl.load_tensors(os.path.join(tl.GOLDEN, f"hvm_{tag}.npz"))
    for c in range(int(e["ncase"])):
        psi = orc.from_packed(t[f"c{c}_psi"])
        op1, op2 = orc.from_packed(t[f"c{c}_op1"]), orc.from_packed(t[f"c{c}_op2"])
        got = orc.to_packed(orc.hvm(psi, op1, op2))
        rel, _ = tl.compare(got, t[f"c{c}_out"])
        assert rel <= 1e-13, (tag, c, rel)
        hl, hr = orc.from_packed(t[f"c{c}_hl"]), orc.from_packed(t[f"c{c}_hr"])
        rel, _ = tl.compare(orc.to_packed(orc.contract(hl, 0, psi, 0)), t[f"c{c}_lpsi"])
        assert rel <= 1e-13
        rel, _ = tl.compare(orc.to_packed(orc.contract(psi, 1, hr, 0)), t[f"c{c}_psir"])
        assert rel <= 1e-13


def test_planted_mistake_is_caught():
    """A wrong sign of the recoupling weight would not pass the hvm check."""
    t, e = tl.load_tensors(os.path.join(tl.GOLDEN, "hvm_su2_real.npz"))
    psi = orc.from_packed(t["c0_psi"])
    op1, op2 = orc.from_packed(t["c0_op1"]), orc.from_packed(t["c0_op2"])
    good = orc.hvm(psi, op1, op2)
    bad = orc.BT(good.sym, good.cplx, 0, good.bonds, {k: -v for k, v in good.blocks.items()})
    rel, _ = tl.compare(orc.to_packed(bad), t["c0_out"])
    assert rel > 0.5


def _state(tens):
    return {k[2:]: v for k, v in tens.items() if k.startswith("S_")}


@pytest.mark.parametrize("name", DMRG_CASES)
def test_oracle_heff_lanczos_svd_on_reference_state(name):
    tens, extra = tl.load_tensors(os.path.join(tl.GOLDEN, f"{name}.npz"))
    ns = int(extra["params"][5])
    il, ir = (int(x) for x in extra["cap_step"])
    got = orc.heff_from_state(_state(tens), extra["hmap"], extra["coup"], ns, il, ir, tens["vec"])
    rel, _ = tl.compare(got, tens["heff"])
    assert rel <= 1e-12, rel
    # Lanczos on the oracle H_eff, flat vectors in block order
    state = {k: orc.from_packed(v) for k, v in _state(tens).items()}
    vec = orc.from_packed(tens["vec"])
    terms = orc.heff_terms(state, np.asarray(extra["hmap"]), extra["coup"], ns, il, ir)
    keys = sorted(vec.blocks)

    def flat(t):
        return np.concatenate([t.blocks[k].ravel(order="F") if k in t.blocks else
                               np.zeros(vec.blocks[k].size, dtype=vec.blocks[k].dtype) for k in keys])

    def unflat(x):
        t = orc.BT(vec.sym, vec.cplx, 0, vec.bonds)
        off = 0
        for k in keys:
            n = vec.blocks[k].size
            t.blocks[k] = x[off:off + n].reshape(vec.blocks[k].shape, order="F")
            off += n
        return t

    H = lambda x: flat(orc.heff(state.get(f"hh_{il}"), state.get(f"hh_{ir}"), terms, unflat(x)))  # noqa: E731
    ndata = sum(v.size for v in vec.blocks.values())
    e, v, nmv = orc.lanczos1(H, flat(vec), min(40, ndata + 1))
    assert abs(e - float(extra["lanczos_e"])) <= 1e-10 * abs(e)
    w, _ = orc.svd_values(orc.from_packed(tens["lanczos_vec"]), int(extra["params"][4]))
    wr = extra["wout"]
    k = min(len(w), len(wr))
    assert np.max(np.abs(np.sort(w[:k]) - np.sort(wr[:k]))) <= 1e-10 * wr.max()


# ---- physics pins --------------------------------------------------------------------
@pytest.mark.parametrize("name,tol", [("dmrg_su2_ring10", 1e-9), ("dmrg_u1_ring10", 1e-9), ("dmrg_u1c_ring10", 1e-9),
                                      ("dmrg_su2_obc12", 1e-9), ("dmrg_su2_j1j2_4x3", 2e-2)])
def test_reference_energies_against_exact_diagonalisation(name, tol):
    """The reference's converged sweep energy versus ED of the same lattice map
    (the J1-J2 cylinder at D=24 is not yet converged: loose bound)."""
    tens, extra = tl.load_tensors(os.path.join(tl.GOLDEN, f"{name}.npz"))
    e_ed = tl.heisenberg_ed(np.asarray(extra["hmap"]), extra["coup"])[0]
    e = float(extra["energies"][-1][-1])
    assert e >= e_ed - 1e-9
    assert abs(e - e_ed) <= tol * abs(e_ed), (e, e_ed)


# ---- oracle vs the reference library itself (where it is built) -----------------------
refbind = pytest.importorskip("refbind")


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built (no reference sources here)")
@pytest.mark.parametrize("name", ["dmrg_su2_ring10", "dmrg_u1_ring10", "dmrg_u1c_ring10"])
def test_oracle_env_updates_against_reference_library(name):
    """Each environment update of the oracle equals the reference's on real
    mid-sweep tensors (both flags, truncated and enlarged bases)."""
    import refbind as rb
    rb.init(physpn=1, max_angm=20, max_dcut=64, psize=1)
    tens, extra = tl.load_tensors(os.path.join(tl.GOLDEN, f"{name}.npz"))
    st = _state(tens)
    checked = 0
    for key, v in st.items():
        if not key.startswith("uu_"):
            continue
        i = int(key.split("_")[1])
        uu = tl.to_ref(v)
        for kind, opkey, flag in [("overlap_transformation", f"hh_{i - 1}", 0),
                                  ("overlap_transformation", f"hh_{i + 1}", 1),
                                  ("operator_transformation", None, 0),
                                  ("operator_transformation", None, 1)]:
            ops = []
            if opkey is not None and opkey in st:
                ops = [st[opkey]]
            elif kind == "operator_transformation":
                nb = i - 1 if flag == 0 else i + 1
                ops = [st[k] for k in st if k.startswith(f"opr_{nb}_")][:3]
            for o in ops:
                b_uu, b_o = orc.from_packed(v), orc.from_packed(o)
                # the reference exits the process on a bond mismatch: check first
                if b_uu.bonds[flag] != orc.inv(b_o.bonds[0]):
                    continue
                if b_uu.bonds[flag] != (b_o.bonds[1] if kind == "overlap_transformation" else b_o.bonds[2]):
                    continue
                ref_out = tl.from_ref(rb.env(kind, uu, uu, tl.to_ref(o), None, flag))
                if int(ref_out["header"][0]) == 0:
                    continue
                got = orc.to_packed(orc.env(kind, b_uu, b_uu, b_o, None, flag))
                rel, _ = tl.compare(got, ref_out)
                assert rel <= 1e-12, (key, kind, flag, rel)
                checked += 1
    assert checked > 0


@pytest.mark.parametrize("name", ["dmrg_su2_ring10", "dmrg_u1_ring10"])
def test_package_lattice_matches_reference_constructor_fixture(name):
    """The package's reference_lattice reproduces the map the reference
    constructor built for the ring fixtures (dmrg_inst.cpp:117-144)."""
    _, extra = tl.load_tensors(os.path.join(tl.GOLDEN, f"{name}.npz"))
    xx, yy = int(extra["params"][2]), int(extra["params"][3])
    got = tl.load_pkg().reference_lattice(xx, yy)
    assert (got == np.asarray(extra["hmap"]).reshape(got.shape)).all()


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built (no reference sources here)")
@pytest.mark.parametrize("xx,yy", [(4, 3), (6, 6), (5, 4), (2, 6)])
def test_package_lattice_matches_reference_constructor(xx, yy):
    d = refbind.RefDMRG(1, 0, xx, yy, 8)
    h, _ = d.couplings()
    got = tl.load_pkg().reference_lattice(xx, yy)
    assert (np.asarray(h).reshape(got.shape) == got).all()


def _ref_sigmas(extra):
    """The reference's local spin operators sigma[0], sigma[1] (dmrg_inst.cpp:222-260),
    read from a reference DMRG of the fixture's symmetry and scalar type."""
    import refbind as rb
    sym, cplx, xx, yy, D = (int(x) for x in extra["params"][:5])
    d = rb.RefDMRG(sym, cplx, xx, yy, D)
    return [tl.from_ref(d.tensor("sigma", 0)), tl.from_ref(d.tensor("sigma", 1))]


def _same_bond(a, b):
    return a[0] == b[0] and list(a[1]) == list(b[1]) and list(a[2]) == list(b[2])


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built (no reference sources here)")
@pytest.mark.parametrize("name", DMRG_CASES)
def test_oracle_initial_and_pairup_updates_against_reference_library(name):
    """overlap_initial, operator_initial and operator_pairup of the oracle equal
    the reference's (block_tensor_inst.cpp:1661-1788, 1282-1469, 1953-2167) on the
    calls prepare_site_operator_from_{left,right}(i, flag) makes
    (dmrg_inst.cpp:440-583): overlap_initial(uu[i], uu[i], flag) with both flags,
    operator_initial(uu[i], uu[i], sigma[flag], flag), and the pair-up of the
    coupling-weighted operator sum sum_k coup[hmap[i][k]] opr[i -+ 1][k] with
    sigma[1 - flag], on every site of the captured mid-sweep state whose
    neighbouring environment is present.  The block Hamiltonian rebuilt from the
    oracle's pieces (overlap_transformation of hh[i -+ 1] plus the pair-up) also
    equals the reference's captured hh[i]."""
    import refbind as rb
    tens, extra = tl.load_tensors(os.path.join(tl.GOLDEN, f"{name}.npz"))
    rb.init(physpn=1, max_angm=20, max_dcut=int(extra["params"][4]), psize=1)
    st = _state(tens)
    ns = int(extra["params"][5])
    hmap = np.asarray(extra["hmap"]).reshape(ns, ns)
    coup = np.asarray(extra["coup"])
    sig_packed = _ref_sigmas(extra)
    sig = [orc.from_packed(s) for s in sig_packed]
    su2 = int(extra["params"][0]) == 1
    il, ir = (int(x) for x in extra["cap_step"])
    counts = {"overlap_initial": 0, "operator_initial": 0, "operator_pairup": 0, "hh": 0}

    def check(kind, uu_p, op_b=None, sg_p=None, flag=0, op_p=None):
        uu_b = orc.from_packed(uu_p)
        got = orc.to_packed(orc.env(kind, uu_b, uu_b, op_b, orc.from_packed(sg_p) if sg_p is not None else None,
                                    flag))
        ref = tl.from_ref(rb.env(kind, tl.to_ref(uu_p), tl.to_ref(uu_p),
                                 tl.to_ref(op_p) if op_p is not None else None,
                                 tl.to_ref(sg_p) if sg_p is not None else None, flag))
        if int(ref["header"][0]) == 0:
            assert int(got["header"][0]) == 0 or not any(got["present"]), (kind, flag)
            return got
        rel, _ = tl.compare(got, ref)
        assert rel <= 1e-12, (name, kind, flag, rel)
        counts[kind] += 1
        return got

    for key, v in st.items():
        if not key.startswith("uu_"):
            continue
        i = int(key.split("_")[1])
        uu_b = orc.from_packed(v)
        for flag in (0, 1):
            check("overlap_initial", v, flag=flag)
            s = sig[flag]
            # operator_initial bond checks (block_tensor_inst.cpp:1296-1311)
            if _same_bond(uu_b.bonds[1 - flag], orc.inv(s.bonds[0])) and _same_bond(uu_b.bonds[1 - flag],
                                                                                     s.bonds[2]):
                check("operator_initial", v, s, None, flag, op_p=sig_packed[flag])
        for flag in (0, 1):
            nb = i - 1 if flag == 0 else i + 1
            # opr[i -+ 1][k] hold block operators of the left (right) environment
            # only on sites up to il (from ir on) in the captured state
            if not (1 <= i <= il if flag == 0 else ir <= i <= ns - 2):
                continue
            ks = range(0, i) if flag == 0 else range(i + 1, ns)
            op = None
            for k in ks:
                o = st.get(f"opr_{nb}_{k}")
                if hmap[i][k] > 0 and o is not None:
                    op = orc.axpy(op, coup[hmap[i][k]], orc.from_packed(o))
            if op is None or op.nbond == 0:
                continue
            s = sig[1 - flag]
            # operator_pairup bond checks (block_tensor_inst.cpp:1963-1988)
            b1 = s.bonds[1] if su2 else orc.inv(s.bonds[1])
            if not (_same_bond(uu_b.bonds[flag], orc.inv(op.bonds[0])) and _same_bond(uu_b.bonds[flag], op.bonds[2])
                    and _same_bond(op.bonds[1], b1)):
                continue
            op_p = orc.to_packed(op)
            pair = check("operator_pairup", v, orc.from_packed(op_p), sig_packed[1 - flag], flag, op_p=op_p)
            # hh[i] = overlap_transformation(uu, uu, hh[i -+ 1], flag) + pair-up (dmrg_inst.cpp:455-458, 471-486)
            hprev = st.get(f"hh_{nb}")
            want = st.get(f"hh_{i}")
            # the sweep built hh[i] in the truncated basis on sites left of the
            # pair (flag 0) and right of it (flag 1); hh[il], hh[ir] are enlarged
            truncated = 1 <= i < il if flag == 0 else ir < i <= ns - 2
            if hprev is None or want is None or not truncated:
                continue
            hp = orc.from_packed(hprev)
            if not (_same_bond(uu_b.bonds[flag], orc.inv(hp.bonds[0])) and _same_bond(uu_b.bonds[flag], hp.bonds[1])):
                continue
            h = orc.env("overlap_transformation", uu_b, uu_b, hp, None, flag)
            h = orc.axpy(h, 1.0, orc.from_packed(pair)) if h.nbond else orc.from_packed(pair)
            rel, _ = tl.compare(orc.to_packed(h), want)
            assert rel <= 1e-12, (name, i, flag, rel)
            counts["hh"] += 1
    print(name, counts)
    assert counts["overlap_initial"] > 0 and counts["operator_initial"] > 0 and counts["operator_pairup"] > 0, counts
    assert counts["hh"] > 0, counts


@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built (no reference sources here)")
def test_oracle_pairup_check_rejects_planted_mistakes():
    """The comparison above is sharp: the pair-up with the transposed spin
    operator (the other flag's sigma) or a sign flip of the recoupling weight
    differs from the reference's output."""
    import refbind as rb
    tens, extra = tl.load_tensors(os.path.join(tl.GOLDEN, "dmrg_su2_obc12.npz"))
    rb.init(physpn=1, max_angm=20, max_dcut=int(extra["params"][4]), psize=1)
    st = _state(tens)
    ns = int(extra["params"][5])
    hmap = np.asarray(extra["hmap"]).reshape(ns, ns)
    coup = np.asarray(extra["coup"])
    sig_packed = _ref_sigmas(extra)
    il = int(extra["cap_step"][0])
    i = il - 1
    op = None
    for k in range(0, i):
        o = st.get(f"opr_{i - 1}_{k}")
        if hmap[i][k] > 0 and o is not None:
            op = orc.axpy(op, coup[hmap[i][k]], orc.from_packed(o))
    assert op is not None
    uu_p = st[f"uu_{i}"]
    uu_b = orc.from_packed(uu_p)
    op_p = orc.to_packed(op)
    ref = tl.from_ref(rb.env("operator_pairup", tl.to_ref(uu_p), tl.to_ref(uu_p), tl.to_ref(op_p),
                             tl.to_ref(sig_packed[1]), 0))
    good = orc.to_packed(orc.env("operator_pairup", uu_b, uu_b, orc.from_packed(op_p),
                                 orc.from_packed(sig_packed[1]), 0))
    assert tl.compare(good, ref)[0] <= 1e-12
    swapped = orc.to_packed(orc.env("operator_pairup", uu_b, uu_b, orc.from_packed(op_p),
                                    orc.from_packed(sig_packed[0]), 0))
    assert tl.compare(swapped, ref)[0] > 1e-3
    orig = orc.fac_pairup
    try:
        orc.fac_pairup = lambda *a, **k: -orig(*a, **k)
        flipped = orc.to_packed(orc.env("operator_pairup", uu_b, uu_b, orc.from_packed(op_p),
                                        orc.from_packed(sig_packed[1]), 0))
    finally:
        orc.fac_pairup = orig
    assert tl.compare(flipped, ref)[0] > 1e-3
