"""Channel-weighted couplings on U(1) drivers: XXZ anisotropy and a z-axis
Dzyaloshinskii-Moriya term (no reference counterpart: the reference's
couplings are one real number per bond, dmrg.hpp:79).  The engine's bond
term becomes coup[c] sum_q w[c][q] O_j(q) O_k(q) over the channel labels
q = -2, 0, +2 of its spin operators.  Exact diagonalisation pins it, and
the exact-diagonalisation helper is itself pinned by Kronecker-built
Hamiltonians written with S^x, S^y, S^z.
"""
import numpy as np
import pytest

import sadmrg_testlib as tl


def _kron_h(ns, bonds):
    """Dense H from S^x, S^y, S^z on every site, bonds = [(i, j, f, Jxy, Jz, D)].
    Each site's basis is (down, up), and the Kronecker product puts site
    ns-1 first, so bit k of a basis index is site k, set = spin up."""
    sx = np.array([[0, 0.5], [0.5, 0]], dtype=complex)
    sy = np.array([[0, 0.5j], [-0.5j, 0]])  # (S^+ - S^-) / 2i with S^+ = |up><down|
    sz = np.array([[-0.5, 0], [0, 0.5]], dtype=complex)

    def op_at(op, i):
        m = np.array([[1.0 + 0j]])
        for k in reversed(range(ns)):
            m = np.kron(m, op if k == i else np.eye(2))
        return m

    H = np.zeros((1 << ns, 1 << ns), dtype=complex)
    for i, j, f, jxy, jz, dm in bonds:
        H += f * (jxy * (op_at(sx, i) @ op_at(sx, j) + op_at(sy, i) @ op_at(sy, j))
                  + jz * op_at(sz, i) @ op_at(sz, j)
                  + dm * (op_at(sx, i) @ op_at(sy, j) - op_at(sy, i) @ op_at(sx, j)))
    return H


@pytest.mark.parametrize("ring", [False, True])
def test_xxz_dm_exact_diagonalisation_brute_force(ring):
    """xxz_dm_hamiltonian equals, element by element, the S^z = 0 block of a
    dense Hamiltonian built from Kronecker products of S^x, S^y, S^z (6
    sites, random Jxy, Jz, D per bond), so the D term's sign and the
    S^+ S^- amplitudes are pinned; the spectra agree too."""
    ns = 6
    rng = np.random.default_rng(5 + ring)
    hmap = np.zeros((ns, ns), dtype=np.int32)
    nb = ns if ring else ns - 1
    coup = np.zeros(nb + 1)
    params, bonds = {}, []
    for b in range(nb):
        i, j = b, (b + 1) % ns
        a, c = min(i, j), max(i, j)
        hmap[a, c] = hmap[c, a] = b + 1
        coup[b + 1] = rng.uniform(0.5, 1.5)
        params[b + 1] = (rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.5), rng.uniform(-0.5, 0.5))
        bonds.append((a, c, coup[b + 1], *params[b + 1]))
    dense = _kron_h(ns, bonds)
    H, states = tl.xxz_dm_hamiltonian(hmap, coup, params)
    block = dense[np.ix_(states, states)]
    assert np.max(np.abs(H.toarray() - block)) <= 1e-14
    # nothing couples S^z = 0 to the other sectors
    rest = [s for s in range(1 << ns) if s not in set(states)]
    assert np.max(np.abs(dense[np.ix_(rest, states)])) == 0
    w = np.sort(np.linalg.eigvalsh(block))
    assert np.allclose(tl.xxz_dm_ed(hmap, coup, params, nlow=4), w[:4], atol=1e-12)


def _driver(sad, ctx, cplx, ns, hmap, coup, D):
    d = sad.DMRG(ctx, sad.U1, cplx, 1, ns, D)
    d.set_couplings(hmap, coup)
    d.set_max_dcut(D)
    return d


def _random_bonds(rng, ns, ring, dm):
    hmap = np.zeros((ns, ns), dtype=np.int32)
    nb = ns if ring else ns - 1
    coup = np.zeros(ns + 1)
    params = {}
    for b in range(nb):
        i, j = b, (b + 1) % ns
        hmap[i, j] = hmap[j, i] = b + 1
        coup[b + 1] = rng.uniform(0.5, 1.5)
        params[b + 1] = (rng.uniform(0.5, 1.5), rng.uniform(0.3, 1.5), rng.uniform(0.0, 0.5) if dm else 0.0)
    return hmap, coup, params


@pytest.mark.gpu
@pytest.mark.parametrize("ring", [False, True])
def test_xxz_real_channel_weights_match_exact_diagonalisation(ctx, sad, ring):
    """Real U(1) driver with per-bond XXZ weights (Jxy != Jz): iDMRG + sweeps
    reach the exact ground energy (1e-9) of the 12-site chain / 10-site ring."""
    ns = 10 if ring else 12
    rng = np.random.default_rng(21 + ring)
    hmap, coup, params = _random_bonds(rng, ns, ring, dm=False)
    e_ed = tl.xxz_dm_ed(hmap, coup, params)[0]
    e_iso = tl.xxz_dm_ed(hmap, coup, {})[0]
    assert abs(e_ed - e_iso) > 1e-3  # the anisotropy matters
    sad.lib().sadmrg_rng_seed(42)
    d = _driver(sad, ctx, 0, ns, hmap, coup, 48)
    d.set_channel_weights(sad.xxz_dm_weights(ns + 1, params))
    d.idmrg()
    for _ in range(4):
        d.sweep()
    assert d.energy() >= e_ed - 1e-9 * abs(e_ed)
    assert abs(d.energy() - e_ed) <= 1e-9 * abs(e_ed), (d.energy(), e_ed)


@pytest.mark.gpu
@pytest.mark.parametrize("ring", [False, True])
def test_dm_complex_channel_weights_match_exact_diagonalisation(ctx, sad, ring):
    """Complex U(1) driver with XXZ + Dzyaloshinskii-Moriya weights: the
    ground energy equals exact diagonalisation of the complex Hamiltonian
    (1e-9).  On the ring the DM twist changes the spectrum beyond what the
    XXZ part gives (the gauge argument that removes it on the open chain
    fails around the ring), which the test checks."""
    ns = 10 if ring else 12
    rng = np.random.default_rng(31 + ring)
    hmap, coup, params = _random_bonds(rng, ns, ring, dm=True)
    e_ed = tl.xxz_dm_ed(hmap, coup, params)[0]
    if ring:
        no_dm = {c: (p[0], p[1], 0.0) for c, p in params.items()}
        assert abs(e_ed - tl.xxz_dm_ed(hmap, coup, no_dm)[0]) > 1e-4
    sad.lib().sadmrg_rng_seed(42)
    d = _driver(sad, ctx, 1, ns, hmap, coup, 48)
    d.set_channel_weights(sad.xxz_dm_weights(ns + 1, params))
    d.idmrg()
    for _ in range(4):
        d.sweep()
    assert d.energy() >= e_ed - 1e-9 * abs(e_ed)
    assert abs(d.energy() - e_ed) <= 1e-9 * abs(e_ed), (d.energy(), e_ed)


@pytest.mark.gpu
def test_channel_weight_errors(ctx, sad):
    """Weights are U(1)-only, must be Hermitian, complex ones need a complex
    driver, and the coupling gradient refuses weighted drivers."""
    ns = 10
    hmap = np.asarray(tl.open_chain(ns))
    coup = np.array([0.0, 1.0] + [0.0] * (ns - 1))
    w = sad.xxz_dm_weights(ns + 1, {1: (1.0, 0.5, 0.2)})
    s = sad.DMRG(ctx, sad.SU2, 0, 1, ns, 16)
    with pytest.raises(sad.SadmrgError, match="U\\(1\\)"):
        s.set_channel_weights(w)
    r = _driver(sad, ctx, 0, ns, hmap, coup, 16)
    with pytest.raises(sad.SadmrgError, match="complex driver"):
        r.set_channel_weights(w)
    bad = w.copy()
    bad[1, 0] = bad[1, 2]  # w(-2) = w(+2) with an imaginary part: not Hermitian
    c = _driver(sad, ctx, 1, ns, hmap, coup, 16)
    with pytest.raises(sad.SadmrgError, match="conjugate"):
        c.set_channel_weights(bad)
    r.set_channel_weights(sad.xxz_dm_weights(ns + 1, {1: (1.0, 0.5, 0.0)}))
    r.idmrg()
    with pytest.raises(sad.SadmrgError, match="isotropic"):
        r.coupling_gradient()


@pytest.mark.gpu
def test_dm_ring_mixed_precision_complex(ctx, sad):
    """Complex XXZ + DM ring with complex64 Krylov matvecs (FP32 MFMA), FP64
    convergence checks and residual refinement: the same ground energy as
    exact diagonalisation (1e-9) and as the FP64 run."""
    ns = 10
    rng = np.random.default_rng(41)
    hmap, coup, params = _random_bonds(rng, ns, True, dm=True)
    e_ed = tl.xxz_dm_ed(hmap, coup, params)[0]
    energies = []
    for prec in (sad.FP64, sad.FP32):
        sad.lib().sadmrg_rng_seed(42)
        d = _driver(sad, ctx, 1, ns, hmap, coup, 48)
        d.set_channel_weights(sad.xxz_dm_weights(ns + 1, params))
        d.set_precision(prec, refine=prec != sad.FP64, refine_tol=1e-7)
        d.idmrg()
        for _ in range(4):
            d.sweep()
        recs = d.records()
        energies.append(d.energy())
        if prec == sad.FP32:
            assert sum(r["matvec32"] for r in recs) > 0  # the complex64 path ran
    for e in energies:
        assert abs(e - e_ed) <= 1e-9 * abs(e_ed), (energies, e_ed)
