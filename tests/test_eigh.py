"""The batched symmetric eigensolver behind the truncation's Gram route
(sadmrg_eigh_batched: persistent tridiagonalisation, bisection, inverse
iteration with Cholesky-QR re-orthonormalisation, back-transformation)
against numpy's LAPACK eigh on the same matrices, with the members of
eigenvalue clusters iterated in parallel lanes and serially
(sadmrg_eigh_clusters).  The matrices are Gram
matrices X X^T whose singular values decay over ten decades, with exact
degenerate pairs (as the entanglement spectra of the symmetric cylinders
have), plus a random symmetric indefinite one; their orders range from 1 to
beyond the largest sector of the D=2048 centre bond (1800: rows longer than
the register-held slices of the reduction), all in one batch, and one with
runs of eigenvalues a few eps apart."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORDERS = [1, 2, 3, 17, 64, 200, 453, 961, 1255, 1800]


def _gram(rng, n, pairs):
    s = np.exp(-np.linspace(0.0, 11.5, n))  # 1 .. 1e-5: eigenvalues 1 .. 1e-10
    if pairs and n > 8:
        s[2:n - 2:7] = s[1:n - 3:7]          # exact degenerate pairs
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * s ** 2) @ q.T


def _buf(sad, ctx, rows, cols, data=None):
    """A device array rows x cols (column-major): the single block of a U(1)
    tensor with one charge sector on each of its two bonds."""
    hdr = np.asarray([2, 0, 1, 1, 0, max(rows, 1), -1, 1, 0, max(cols, 1)], dtype=np.int32)
    x = np.zeros(max(rows, 1) * max(cols, 1))
    if data is not None:
        x[:data.size] = np.asarray(data, dtype=np.float64).ravel(order="F")
    return sad.Tensor.from_packed(ctx, sad.Packed(0, 0, hdr, np.ones(1, np.uint8), x))


def _get(t, rows, cols):
    return t.packed().data[:rows * cols].reshape((rows, cols), order="F")


def _solve(sad, ctx, mats, ks):
    n = [m.shape[0] for m in mats]
    da = [_buf(sad, ctx, x, x, m) for x, m in zip(n, mats)]
    dw = [_buf(sad, ctx, x, 1) for x in n]
    dz = [_buf(sad, ctx, x, max(k, 1)) for x, k in zip(n, ks)]
    sad.eigh_batched(ctx, n, [t.device_ptr() for t in da], [max(x, 1) for x in n], [t.device_ptr() for t in dw],
                     ks, [t.device_ptr() for t in dz], [max(x, 1) for x in n])
    w = [_get(t, x, 1)[:, 0] for t, x in zip(dw, n)]
    z = [_get(t, x, max(k, 1))[:, :k] for t, x, k in zip(dz, n, ks)]
    return w, z


def _near_ties(rng, n):
    """Runs of eigenvalues a few eps apart (clusters inverse iteration
    cannot separate, but not exactly degenerate) between well separated
    ones."""
    lam = np.linspace(1.0, 0.01, n)
    eps = np.finfo(float).eps
    for c in range(0, n - 8, 16):
        lam[c:c + 6] = lam[c] * (1.0 - eps * np.array([0, 2, 3, 7, 8, 20]))
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * lam) @ q.T


@pytest.fixture(params=[0, 1], ids=["parallel_members", "serial_clusters"])
def cluster_mode(sad, request):
    prev = sad.eigh_clusters(request.param)[0]
    yield request.param
    sad.eigh_clusters(prev)


def test_eigh_batched_matches_lapack(sad, ctx, cluster_mode):
    rng = np.random.default_rng(2024)
    mats = [_gram(rng, x, pairs=True) for x in ORDERS]
    b = rng.standard_normal((300, 300))
    mats.append(0.5 * (b + b.T))
    mats.append(_near_ties(rng, 240))
    ks = [max(0, (2 * m.shape[0]) // 5) for m in mats]
    ks[0] = 1
    w, z = _solve(sad, ctx, mats, ks)
    for m, wq, zq, k in zip(mats, w, z, ks):
        n = m.shape[0]
        ref, vref = np.linalg.eigh(m)
        scale = np.abs(ref).max()
        # eigenvalues: absolute accuracy of a backward-stable solver
        assert np.max(np.abs(wq - ref)) <= 1e-13 * n ** 0.5 * scale + 1e-300, (n, np.max(np.abs(wq - ref)))
        if k == 0:
            continue
        lam = wq[::-1][:k]
        # residuals and orthonormality of the eigenvectors
        res = np.linalg.norm(m @ zq - zq * lam, axis=0)
        assert res.max() <= 1e-12 * scale, (n, res.max())
        assert np.max(np.abs(zq.T @ zq - np.eye(k))) <= 1e-13, (n, np.max(np.abs(zq.T @ zq - np.eye(k))))
        # the kept invariant subspace equals LAPACK's where the cut has a gap
        gap = ref[n - k] - ref[n - k - 1] if k < n else 1.0
        if gap > 1e-8 * scale:
            vk = vref[:, ::-1][:, :k]
            sv = np.linalg.svd(vk.T @ zq, compute_uv=False)
            assert sv.min() >= 1 - 1e-10, (n, sv.min())


def test_eigh_batched_validates(sad, ctx):
    a = _buf(sad, ctx, 4, 4)
    w = _buf(sad, ctx, 4, 1)
    with pytest.raises(sad.SadmrgError):
        sad.eigh_batched(ctx, [4], [a.device_ptr()], [3], [w.device_ptr()])   # lda < n
    with pytest.raises(sad.SadmrgError):
        sad.eigh_batched(ctx, [4], [a.device_ptr()], [4], [w.device_ptr()], [5], [a.device_ptr()], [4])  # k > n
    with pytest.raises(sad.SadmrgError):
        sad.eigh_batched(ctx, [5000], [a.device_ptr()], [5000], [w.device_ptr()])  # order beyond the solver
    with pytest.raises(sad.SadmrgError):
        sad.eigh_clusters(2)
