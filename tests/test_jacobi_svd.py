"""Small-sector truncation SVD by the one-sided Jacobi kernel
(sadmrg_svd_jacobi; replaces gesdd, the stand-in for the reference's gesvd
of dense_inst.cpp:486, on real sectors with min(m, n) <= 64).

On the centre bond of the U(1) open chain L = 32 at m = 64 (the C1
configuration) the sectors go through the kernel:
  * every singular value the kept rule can select (S / S_max > 1e-10 in its
    sector, dense_inst.cpp:497-502) equals numpy's SVD of the block to 1e-12
    of itself plus 1e-13 of the block's largest (numpy's gesdd has an
    absolute error of a few eps S_max), against 1e-10 S_max for the gesdd
    and Gram routes;
  * psi rebuilt from the new site tensors (prepare_input_vector) equals psi
    when the cap does not bind, and is the orthogonal projection onto the
    kept states when it binds, so U and V' are orthonormal and match S;
  * gesdd (the kernel switched off) gives the same values to 1e-10 S_max.
Both sweep directions' absorption of S are covered ((p1, p2) = (1, 0) and
(0, 1)); iDMRG steps ((0, 0)) stay on gesdd, which the counters confirm."""
import numpy as np
import pytest

import sadmrg_testlib as tl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chain_state(ctx, sad):
    L, m = 32, 64
    coup = np.zeros(L + 1)
    coup[1] = 1.0
    sad.lib().sadmrg_rng_seed(42)
    d = sad.DMRG(ctx, sad.U1, 0, 1, L, m)
    d.set_couplings(np.asarray(sad.open_chain(L), dtype=np.int32), coup)
    d.set_max_dcut(m)
    d.idmrg()
    d.sweep()
    return d, L


@pytest.fixture
def jacobi_default(sad):
    prev = sad.svd_jacobi()[0]
    yield
    sad.svd_jacobi(prev)


@pytest.mark.parametrize("p1,p2", [(1.0, 0.0), (0.0, 1.0), (0.0, 0.0)])  # (0, 0): gesdd
def test_jacobi_truncation(ctx, sad, chain_state, jacobi_default, p1, p2):
    d, L = chain_state
    il, ir = L // 2 - 1, L // 2
    vec = d.input_vector(il, ir)
    v = tl.from_dev(vec)
    blocks = tl.blocks(v)
    def fits(b):  # the kernel's limits: min(m, n) <= 64, working set in 64 KB of LDS
        q, p_ = min(b.shape), max(b.shape)
        qe = q + (q & 1)
        return q <= 64 and p_ * qe + qe * qe <= 8192
    njac = sum(fits(b) for b in blocks.values())
    assert njac >= len(blocks) - 2 and njac > 0, [b.shape for b in blocks.values()]
    per = [np.linalg.svd(b, compute_uv=False) for b in blocks.values()]
    smax = max(x[0] for x in per)
    sel = [x[x > 1e-10 * x[0]] if x[0] > 0 else x[:1] for x in per]
    cand = np.concatenate(sel)
    tol = np.concatenate([1e-12 * x + 1e-13 * x[0] + 1e-300 for x in sel])  # zero blocks: exact zeros
    order = np.argsort(-cand, kind="stable")
    cand, tol = cand[order], tol[order]

    sad.svd_jacobi(64)
    _, n0, r0 = sad.svd_jacobi()
    c = d.clone()
    c.set_max_dcut(100000)
    w = np.sort(np.asarray(c.svd(vec, il, ir, p1, p2)))[::-1]
    _, n1, r1 = sad.svd_jacobi()
    assert n1 - n0 == (njac if p1 + p2 > 0 else 0) and r1 == r0, (n0, n1, r0, r1, njac)
    assert len(w) == len(cand), (len(w), len(cand))
    if p1 + p2 == 0:  # gesdd: absolute error up to ~eps S_max^2 / S
        tol = 1e-10 * smax + 4 * np.finfo(float).eps * smax ** 2 / np.maximum(cand, 1e-300)
    err = np.abs(w - cand) / tol
    assert np.max(err) <= 1.0, (float(np.max(err)), float(cand[np.argmax(err)] / smax))
    # cap binding: the kept values
    keep = len(cand) // 2
    ck = d.clone()
    ck.set_max_dcut(keep)
    wk = np.asarray(ck.svd(vec, il, ir, p1, p2))
    assert np.max(np.abs(np.sort(wk[:keep])[::-1] - cand[:keep]) / tol[:keep]) <= 1.0
    if p1 + p2 > 0:
        # sweep steps (one side absorbs S): psi rebuilt from the site
        # tensors is psi itself, or its orthogonal projection onto the kept
        # states when the cap binds
        back = tl.from_dev(c.input_vector(il, ir))
        rel_b, _ = tl.compare(back, v)
        assert rel_b <= 1e-12, rel_b
        back = tl.blocks(tl.from_dev(ck.input_vector(il, ir)))
        dot = sum(float(np.sum(back[k] * blocks[k])) for k in back if k in blocks)
        nrm = sum(float(np.sum(x * x)) for x in back.values())
        target = float(np.sum(cand[:keep] ** 2))
        assert abs(nrm - target) <= 1e-12 * target and abs(dot - target) <= 1e-12 * target, (nrm, dot, target)

    # gesdd on the same sectors
    sad.svd_jacobi(0)
    c = d.clone()
    c.set_max_dcut(100000)
    wg = np.sort(np.asarray(c.svd(vec, il, ir, p1, p2)))[::-1]
    m = min(len(wg), len(w))
    assert np.max(np.abs(wg[:m] - w[:m])) <= 1e-10 * smax


def test_jacobi_switch_limits(sad, jacobi_default):
    with pytest.raises(sad.SadmrgError):
        sad.svd_jacobi(89)
    assert sad.svd_jacobi(88)[0] == 64
    assert sad.svd_jacobi(-1)[0] == 88
