"""e8qlib.prep (PCA and IVF assignment for the ME8Q path) against an fp64
numpy restatement of the reference recipe (datasets/prepare.py compute_pca
:567-615, assign_ivf :537-560), on the CPU and on the GPU.

Tolerances: the fp64 mean, covariance and eigenvalues within 1e-10 of their
scale (only the summation order differs); eigenvectors equal up to sign
(LAPACK fixes their sign from the matrix, and the two covariances differ in
the last bits only); the float32 transform within 2e-6 of the largest
coordinate after aligning the signs; the assignment equal wherever the two
nearest centroids are not within 1e-5 relative of each other.
"""
import numpy as np
import pytest

from e8qlib import prep


def _reference_pca(x):
    x64 = x.astype(np.float64)
    n = x.shape[0]
    mu = x64.sum(0) / n
    cov = (x64.T @ x64) / n - np.outer(mu, mu)
    evals, evecs = np.linalg.eigh(cov)
    return mu, evecs[:, ::-1], evals[::-1].clip(min=0)


def _reference_assign(x, cent):
    c2 = (cent.astype(np.float64) ** 2).sum(1)
    d = (x.astype(np.float64) ** 2).sum(1)[:, None] - 2.0 * (x @ cent.T) + c2[None, :]
    return d.argmin(1).astype(np.int32), d


def _data(seed=5, n=6000, d=48, nc=40):
    rng = np.random.default_rng(seed)
    lam = np.arange(1, d + 1, dtype=np.float64) ** -1.2
    rot, _ = np.linalg.qr(rng.standard_normal((d, d)))
    x = ((rng.standard_normal((n, d)) * np.sqrt(lam)) @ rot.T + 0.3).astype(np.float32)
    cent = x[rng.choice(n, nc, replace=False)].copy()
    return x, cent


def _check(device):
    x, cent = _data()
    p = prep.pca_fit(x, block=1000, device=device)
    mu, V, ev = _reference_pca(x)
    np.testing.assert_allclose(p.mean, mu, rtol=0, atol=1e-12 * np.abs(mu).max())
    np.testing.assert_allclose(p.variance, ev, rtol=0, atol=1e-10 * ev.max())
    assert np.all(np.diff(p.variance) <= 0)
    dots = np.abs((p.components * V).sum(0))
    np.testing.assert_allclose(dots, 1.0, atol=1e-8)
    sign = np.sign((p.components * V).sum(0))
    y = prep.pca_transform(p, x, block=1000, device=device).cpu().numpy()
    ref_y = ((x.astype(np.float64) - mu) @ (V * sign)).astype(np.float32)
    np.testing.assert_allclose(y, ref_y, rtol=0, atol=2e-6 * np.abs(ref_y).max())
    ids = prep.assign_ivf(x, cent, block=1000, device=device).cpu().numpy()
    rids, d = _reference_assign(x, cent)
    part = np.partition(d, 1, axis=1)
    clear = (part[:, 1] - part[:, 0]) > 1e-5 * np.abs(part[:, 0]).max()
    assert clear.mean() > 0.99
    np.testing.assert_array_equal(ids[clear], rids[clear])


def test_prep_cpu():
    _check("cpu")


def test_prep_rejects_empty():
    with pytest.raises(ValueError):
        prep.pca_fit(np.zeros((0, 8), np.float32), device="cpu")


@pytest.mark.gpu
def test_prep_gpu():
    _check("cuda:0")
