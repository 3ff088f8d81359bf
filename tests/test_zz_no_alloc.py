"""The allocation-free eigenvalue entry point (sadmrg_eigvalsh_batched_noalloc,
static per-device scratch) against the allocating one it mirrors
(sadmrg_eigh_batched with k = 0) on the same matrices, over two calls in a
row -- the second batch smaller than the first, so stale scratch contents
from the first call would show -- and against numpy's LAPACK eigvalsh.
Also: the context's device memory is unchanged by the call, and a batch
larger than the scratch is rejected before any work."""
import numpy as np
import pytest

from test_eigh import _buf, _get, _gram

pytestmark = pytest.mark.gpu


def _mats(rng, orders):
    mats = [_gram(rng, x, pairs=True) for x in orders]
    b = rng.standard_normal((150, 150))
    mats.append(0.5 * (b + b.T))  # indefinite
    return mats


def _eigvals(sad, ctx, mats, noalloc):
    n = [m.shape[0] for m in mats]
    da = [_buf(sad, ctx, x, x, m) for x, m in zip(n, mats)]
    dw = [_buf(sad, ctx, x, 1) for x in n]
    ptr_a = [t.device_ptr() for t in da]
    ptr_w = [t.device_ptr() for t in dw]
    lda = [max(x, 1) for x in n]
    if noalloc:
        before = ctx.memory()
        sad.eigvalsh_batched_noalloc(ctx, n, ptr_a, lda, ptr_w)
        ctx.sync()
        assert ctx.memory() == before, (before, ctx.memory())
    else:
        sad.eigh_batched(ctx, n, ptr_a, lda, ptr_w)
    return [_get(t, x, 1)[:, 0] for t, x in zip(dw, n)]


@pytest.mark.parametrize("orders", [[1, 3, 64, 453, 961, 1255], [2, 17, 200]])
def test_noalloc_eigenvalues_match_allocating_entry_point(sad, ctx, orders):
    rng = np.random.default_rng(len(orders))
    mats = _mats(rng, orders)
    for call in range(2):  # the second call reuses the scratch the first left behind
        got = _eigvals(sad, ctx, mats if call == 0 else mats[1:], noalloc=True)
        want = _eigvals(sad, ctx, mats if call == 0 else mats[1:], noalloc=False)
        for m, g, wv in zip(mats if call == 0 else mats[1:], got, want):
            n = m.shape[0]
            ref = np.linalg.eigvalsh(m)
            scale = np.abs(ref).max()
            # same kernels, same launch plan: the same values
            assert np.max(np.abs(g - wv)) <= 1e-15 * scale, (call, n, np.max(np.abs(g - wv)))
            assert np.max(np.abs(g - ref)) <= 1e-13 * n ** 0.5 * scale, (call, n, np.max(np.abs(g - ref)))


def test_noalloc_rejects_batches_beyond_its_scratch(sad, ctx):
    a = _buf(sad, ctx, 2048, 2048)
    w = _buf(sad, ctx, 2048, 1)
    cnt = 12  # 12 * 2048 rows * 88 bytes > 2 MiB
    with pytest.raises(sad.SadmrgError) as e:
        sad.eigvalsh_batched_noalloc(ctx, [2048] * cnt, [a.device_ptr()] * cnt, [2048] * cnt, [w.device_ptr()] * cnt)
    assert e.value.code == 1
    with pytest.raises(sad.SadmrgError) as e:
        sad.eigvalsh_batched_noalloc(ctx, [4], [a.device_ptr()], [3], [w.device_ptr()])  # lda < n
    assert e.value.code == 1
