"""Image-quality measures and the convergence rule of the benchmark.

RMSE in HU: 1000 * RMSE(mu) / mu_water with mu_water = 0.019 /mm, the HU map
mu = (HU + 1000) / 1000 * 0.019 of ReconNet4D_no_tf/ProjectorTester.cu:27-28,
taken over the circular field of view.
"""

import numpy as np

from .geometry import HU_WATER

__all__ = ["rmse_hu", "convergence_pass"]


def rmse_hu(x, truth, mask):
    """Per-slice RMSE in HU of images [B, nx, ny] (numpy or torch) over `mask`."""
    d = (x - truth) * mask
    n = mask.sum()
    return (1000.0 / HU_WATER) * ((d * d).sum(axis=(1, 2)) / n) ** 0.5 if isinstance(d, np.ndarray) else \
        (1000.0 / HU_WATER) * ((d * d).sum(dim=(1, 2)) / n).sqrt()


def convergence_pass(mean_rmse, rel_update, rtol=0.01, utol=1e-4):
    """First pass k (1-based) with mean_rmse[k] <= (1 + rtol) * mean_rmse[-1]
    and rel_update[k] < utol; len(mean_rmse) (k_max) if no pass qualifies."""
    mean_rmse = np.asarray(mean_rmse, np.float64)
    rel_update = np.asarray(rel_update, np.float64)
    if mean_rmse.size == 0 or mean_rmse.shape != rel_update.shape:
        raise ValueError("need one RMSE and one update per pass")
    ok = (mean_rmse <= (1.0 + rtol) * mean_rmse[-1]) & (rel_update < utol)
    idx = np.flatnonzero(ok)
    return int(idx[0]) + 1 if idx.size else int(mean_rmse.size)
