"""Block-tensor vector algebra through the C ABI against the reference library:
inner_prod_c, operator+= / -=, *=, normalize_vector, operator=(double),
take_conjugate and shift (block_tensor.hpp:106-122, block_tensor_inst.cpp:197-376,
685-702, 914-948), on seeded random tensors of both symmetries and scalar
types, including blocks present in only one operand.  Also the long-basis
Lanczos solve (host-driven re-orthogonalisation over 64-column panels)
against a CPU Krylov projection with the oracle's H_eff."""
import ctypes
import os

import numpy as np
import pytest

import sadmrg_testlib as tl
import sadmrg_oracle as orc

pytestmark = pytest.mark.gpu

KINDS = [(1, 0), (1, 1), (0, 0), (0, 1)]  # (sym, cplx): SU(2)/U(1) x real/complex


def _bonds(sym, rank):
    if sym == 1:
        a = (1, [0, 2, 4], [5, 7, 3])
        b = (-1, [0, 2, 4], [4, 6, 2])
        c = (1, [0, 2], [2, 1])
    else:
        a = (1, [-2, 0, 2], [5, 7, 3])
        b = (-1, [-2, 0, 2], [4, 6, 2])
        c = (1, [-2, 0, 2], [1, 2, 1])
    return [a, b] if rank == 2 else [a, c, b]


def _refbind():
    import refbind as rb
    if not rb.available():
        pytest.skip("oracle/_ref not built")
    rb.init(physpn=1, max_angm=20, max_dcut=64, psize=1)
    return rb


def _ref_op(rb, t, op, a=0.0, b=0.0):
    rb.lib().ref_tensor_op(t.ptr, op, a, b)
    return t


@pytest.mark.parametrize("sym,cplx", KINDS)
def test_dot_and_axpy_match_reference(ctx, sad, sym, cplx):
    rb = _refbind()
    rng = np.random.default_rng(17 + 2 * sym + cplx)
    for rank in (2, 3):
        bonds = _bonds(sym, rank)
        locspin = 0
        x = tl.random_tensor(rng, sym, cplx, locspin, bonds)
        y = tl.random_tensor(rng, sym, cplx, locspin, bonds, density=0.6)
        # inner_prod_c: x^H y over blocks present in both
        got = tl.to_dev(ctx, x).dot(tl.to_dev(ctx, y))
        o = np.zeros(2)
        rx, ry = tl.to_ref(x), tl.to_ref(y)  # the handles own the reference tensors: keep them alive
        rb.lib().ref_inner_prod_c(rx.ptr, ry.ptr, o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        want = complex(o[0], o[1])
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (rank, got, want)
        # y += x and y -= x; y null where x is present gives a copy
        for sign, code in ((1.0, 0), (-1.0, 1)):
            ty = tl.to_dev(ctx, y).axpy(sign, tl.to_dev(ctx, x))
            ry = tl.to_ref(y)
            rb.lib().ref_tensor_binop(ry.ptr, code, rx.ptr)
            rel, _ = tl.compare(tl.from_dev(ty), tl.from_ref(ry))
            assert rel <= 1e-15, (rank, sign, rel)
        # y += (a + ib) x: the reference scales a copy of x, then adds
        a, b = 0.75, (-0.4 if cplx else 0.0)
        ty = tl.to_dev(ctx, y).axpy(a, tl.to_dev(ctx, x), im=b)
        rx = tl.to_ref(x)
        _ref_op(rb, rx, 9 if cplx else 4, a, b)
        ry = tl.to_ref(y)
        rb.lib().ref_tensor_binop(ry.ptr, 0, rx.ptr)
        rel, _ = tl.compare(tl.from_dev(ty), tl.from_ref(ry))
        assert rel <= 1e-15, (rank, rel)


@pytest.mark.parametrize("sym,cplx", KINDS)
def test_transforms_match_reference(ctx, sad, sym, cplx):
    """*=, normalize_vector, operator=(double), take_conjugate(bond), shift."""
    rb = _refbind()
    rng = np.random.default_rng(29 + 2 * sym + cplx)
    for rank in (2, 3):
        bonds = _bonds(sym, rank)
        x = tl.random_tensor(rng, sym, cplx, 0, bonds, density=0.8)
        cases = [(4, 1.7, 0.0, 4), (5, 0.0, 0.0, 5), (8, 0.3, 0.0, 8)]
        if cplx:
            cases.append((4, 0.6, -1.1, 9))  # engine op 4 is *= (a + ib)
        cases += [(2, float(k), 0.0, 2) for k in range(rank)]
        if rank == 3:
            cases += [(7, 0.0, 1.0, 7), (7, 0.0, 2.0, 7), (7, 2.0, 0.0, 7), (7, 1.0, 2.0, 7)]
        for op, a, b, ref_op in cases:
            got = tl.from_dev(tl.to_dev(ctx, x).transform(op, a, b))
            want = tl.from_ref(_ref_op(rb, tl.to_ref(x), ref_op, a, b))
            rel, _ = tl.compare(got, want)
            assert rel <= 1e-14, (rank, op, a, b, rel)


def test_dot_rejects_mismatched_structures(ctx, sad):
    """Two tensors whose present blocks differ in size: kErrInvalid before any
    kernel runs (the reference exits, dense_inst.cpp:637)."""
    rng = np.random.default_rng(5)
    b1 = _bonds(1, 2)
    b2 = [(1, [0, 2, 4], [5, 8, 3]), (-1, [0, 2, 4], [4, 6, 2])]
    x = tl.to_dev(ctx, tl.random_tensor(rng, 1, 0, 0, b1))
    y = tl.to_dev(ctx, tl.random_tensor(rng, 1, 0, 0, b2))
    with pytest.raises(sad.SadmrgError) as e:
        x.dot(y)
    assert e.value.code == 1
    z = tl.to_dev(ctx, tl.random_tensor(rng, 1, 1, 0, b1))
    with pytest.raises(sad.SadmrgError):
        x.dot(z)


def _state_plan(ctx, sad, cgc, name):
    tens, extra = tl.load_tensors(os.path.join(tl.GOLDEN, f"{name}.npz"))
    ns = int(extra["params"][5])
    il, ir = (int(x) for x in extra["cap_step"])
    st = {k[2:]: orc.from_packed(v) for k, v in tens.items() if k.startswith("S_")}
    terms = orc.heff_terms(st, np.asarray(extra["hmap"]), extra["coup"], ns, il, ir)
    cg2 = 1 if tens["vec"]["sym"] == 1 else 0
    psi = tl.to_dev(ctx, tens["vec"], cg2)
    hl = tl.to_dev(ctx, orc.to_packed(st[f"hh_{il}"]))
    hr = tl.to_dev(ctx, orc.to_packed(st[f"hh_{ir}"]))
    o1 = [tl.to_dev(ctx, orc.to_packed(a)) for a, _ in terms]
    o2 = [tl.to_dev(ctx, orc.to_packed(b)) for _, b in terms]
    plan = sad.Heff(ctx, cgc, hl, hr, o1, o2, psi)
    return plan, psi, st, terms, tens, extra, il, ir


def test_lanczos_long_basis_matches_krylov_projection(ctx, sad, cgc):
    """An 80-vector basis with no intermediate convergence check: the
    re-orthogonalisation runs over two 64-column panels.  The energy equals
    the lowest eigenvalue of H_eff projected on the same 80-dimensional
    Krylov space, built on the CPU with the oracle's H_eff and twice-repeated
    full Gram-Schmidt."""
    plan, psi, st, terms, tens, extra, il, ir = _state_plan(ctx, sad, cgc, "dmrg_u1_obc32")
    ml = 80
    e, info = plan.lanczos_ex(psi, mlanc=ml, check_every=1000, device_steps=0)
    assert info["steps"] == ml and info["matvec64"] >= ml, info
    vec = orc.from_packed(tens["vec"])
    keys = sorted(vec.blocks)

    def flat(t):
        return np.concatenate([t.blocks[k].ravel(order="F") if k in t.blocks else
                               np.zeros(vec.blocks[k].size) for k in keys])

    def unflat(x):
        t = orc.BT(vec.sym, vec.cplx, 0, vec.bonds)
        off = 0
        for k in keys:
            n = vec.blocks[k].size
            t.blocks[k] = x[off:off + n].reshape(vec.blocks[k].shape, order="F")
            off += n
        return t

    def H(x):
        return flat(orc.heff(st.get(f"hh_{il}"), st.get(f"hh_{ir}"), terms, unflat(x)))

    q = flat(vec)
    Q = np.zeros((q.size, ml))
    Q[:, 0] = q / np.linalg.norm(q)
    for m in range(1, ml):
        w = H(Q[:, m - 1])
        for _ in range(2):
            w -= Q[:, :m] @ (Q[:, :m].T @ w)
        Q[:, m] = w / np.linalg.norm(w)
    T = Q.T @ np.column_stack([H(Q[:, j]) for j in range(ml)])
    e_krylov = np.linalg.eigvalsh(0.5 * (T + T.T))[0]
    assert abs(e - e_krylov) <= 1e-10 * abs(e_krylov), (e, e_krylov)


def test_lanczos_options_are_validated(ctx, sad, cgc):
    plan, psi, *_ = _state_plan(ctx, sad, cgc, "dmrg_su2_ring10")
    for bad in ({"mlanc": 0}, {"mlanc": 2048}, {"check_every": 0}, {"refine_krylov": 0}, {"refine_tol": 0.0}):
        with pytest.raises(sad.SadmrgError) as e:
            plan.lanczos_ex(psi, **bad)
        assert e.value.code == 1, bad
    with pytest.raises(sad.SadmrgError):
        plan.lanczos(psi, mlanc=5000)
