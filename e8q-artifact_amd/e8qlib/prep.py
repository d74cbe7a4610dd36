"""Preprocessing for the ME8Q path at scale: PCA to the variance-ordered
domain ME8Q consumes, and IVF list assignment.

The reference's recipe (datasets/prepare.py: compute_pca :567-615,
assign_ivf :537-560) runs on fvecs files with numpy; these functions run the
same arithmetic on arrays, with the row-block GEMMs on the GPU:

* pca_fit: mean and covariance accumulated in fp64 over row blocks,
  cov = X^T X / n - mu mu^T; eigen-decomposition on the host (numpy eigh,
  fp64); eigenvalues in descending order, clipped at 0 (the per-dimension
  variance Me8qIndex.build takes), eigenvectors in the same order (column i
  is the i-th component).
* pca_transform: y = (x - mu) V in fp64, stored as float32.
* assign_ivf: argmin over the centroids of ||x||^2 - 2 x.c + ||c||^2, the
  norms in fp64 and x.c a float32 product, as the reference's numpy path
  computes it; ties go to the lower centroid id.

Arrays may be numpy arrays or torch tensors.  The work runs on `device`
(default "cuda:0"; a torch tensor input runs where it lives); "cpu" must be
asked for explicitly.
"""
from dataclasses import dataclass

import numpy as np

BLOCK = 65536   # rows per block (the reference's RECORD_BLOCK)


def _torch():
    import torch
    return torch


def _device(x, device):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.device
    return torch.device(device if device is not None else "cuda:0")


def _rows(x, s, e, dev, dtype):
    """Rows s..e of x as a tensor of `dtype` on `dev`."""
    torch = _torch()
    b = x[s:e]
    if not isinstance(b, torch.Tensor):
        b = torch.from_numpy(np.ascontiguousarray(b))
    return b.to(device=dev, dtype=dtype)


@dataclass
class PCA:
    mean: np.ndarray        # fp64 [d]
    components: np.ndarray  # fp64 [d, d], column i = i-th eigenvector (descending eigenvalue)
    variance: np.ndarray    # fp64 [d], descending, >= 0


def pca_fit(x, block=BLOCK, device=None):
    """PCA of the rows of x (n x d) as compute_pca fits it."""
    torch = _torch()
    n, d = int(x.shape[0]), int(x.shape[1])
    if n == 0:
        raise ValueError("pca_fit: no rows")
    dev = _device(x, device)
    mu = torch.zeros(d, dtype=torch.float64, device=dev)
    cov = torch.zeros(d, d, dtype=torch.float64, device=dev)
    for s in range(0, n, block):
        xb = _rows(x, s, min(n, s + block), dev, torch.float64)
        mu += xb.sum(0)
        cov += xb.T @ xb
    mu /= n
    cov = cov / n - torch.outer(mu, mu)
    evals, evecs = np.linalg.eigh(cov.cpu().numpy())   # ascending
    evals = evals[::-1].clip(min=0)
    evecs = evecs[:, ::-1]
    return PCA(mean=mu.cpu().numpy(), components=np.ascontiguousarray(evecs), variance=np.ascontiguousarray(evals))


def pca_transform(pca, x, block=BLOCK, device=None, out=None):
    """(x - mean) @ components in fp64, as float32.  Returns a tensor on the
    work device (written into `out` when given, which may alias x)."""
    torch = _torch()
    n, d = int(x.shape[0]), int(x.shape[1])
    dev = _device(x, device)
    mu = torch.from_numpy(pca.mean).to(dev)
    V = torch.from_numpy(pca.components).to(dev)
    if out is None:
        out = torch.empty((n, V.shape[1]), dtype=torch.float32, device=dev)
    for s in range(0, n, block):
        e = min(n, s + block)
        out[s:e] = ((_rows(x, s, e, dev, torch.float64) - mu) @ V).to(torch.float32)
    return out


def assign_ivf(x, centroids, block=BLOCK, device=None):
    """Nearest centroid of every row of x: int32 ids on the work device."""
    torch = _torch()
    n = int(x.shape[0])
    dev = _device(x, device)
    c = centroids if isinstance(centroids, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(centroids))
    c32 = c.to(device=dev, dtype=torch.float32)
    c2 = (c32.to(torch.float64) ** 2).sum(1)
    ids = torch.empty(n, dtype=torch.int32, device=dev)
    for s in range(0, n, block):
        e = min(n, s + block)
        xb = _rows(x, s, e, dev, torch.float32)
        d = (xb.to(torch.float64) ** 2).sum(1)[:, None] - 2.0 * (xb @ c32.T).to(torch.float64) + c2[None, :]
        ids[s:e] = d.argmin(1).to(torch.int32)
    return ids
