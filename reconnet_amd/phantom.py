"""Seeded random-ellipse phantoms and transmission noise.

Phantom recipe: numpy ``default_rng(seed)``; per ellipse a centre at radius
110 sqrt(U) mm and angle 2 pi U, semi-axes U[5, 80] mm, rotation U[0, pi),
attenuation U[0, 0.04] /mm; overlapping ellipses add; 4 x 4 supersampled
rasterisation on the image grid (pixel centres at (i - (n-1)/2) d).
"""

import numpy as np

__all__ = ["ellipse_params", "rasterize", "phantom", "phantom_batch", "add_poisson_noise", "ellipsoid_params",
           "phantom3d"]


def ellipse_params(seed, n_ellipses=10, r_max=110.0, axis_range=(5.0, 80.0), mu_max=0.04):
    rng = np.random.default_rng(seed)
    out = np.zeros((n_ellipses, 6), np.float64)
    for k in range(n_ellipses):
        r = r_max * np.sqrt(rng.uniform())
        phi = 2 * np.pi * rng.uniform()
        a = rng.uniform(*axis_range)
        b = rng.uniform(*axis_range)
        rot = rng.uniform(0, np.pi)
        mu = rng.uniform(0, mu_max)
        out[k] = [r * np.cos(phi), r * np.sin(phi), a, b, rot, mu]
    return out


def rasterize(params, nx, ny, dx, dy, supersample=4, device=None):
    """Rasterise ellipses; returns float32 [nx, ny] (numpy, or torch on `device`)."""
    if device is not None:
        import torch

        dt = torch.float64
        off = (torch.arange(supersample, dtype=dt, device=device) + 0.5) / supersample - 0.5
        xs = ((torch.arange(nx, dtype=dt, device=device) - (nx - 1) / 2)[:, None] + off[None, :]).reshape(-1) * dx
        ys = ((torch.arange(ny, dtype=dt, device=device) - (ny - 1) / 2)[:, None] + off[None, :]).reshape(-1) * dy
        X, Y = xs[:, None], ys[None, :]
        img = torch.zeros((nx * supersample, ny * supersample), dtype=dt, device=device)
        for cx, cy, a, b, rot, mu in params:
            c, s = np.cos(rot), np.sin(rot)
            u = (X - cx) * c + (Y - cy) * s
            v = -(X - cx) * s + (Y - cy) * c
            img += mu * ((u / a) ** 2 + (v / b) ** 2 <= 1.0).to(dt)
        img = img.reshape(nx, supersample, ny, supersample).mean(dim=(1, 3))
        return img.to(torch.float32)
    off = (np.arange(supersample) + 0.5) / supersample - 0.5
    xs = (((np.arange(nx) - (nx - 1) / 2)[:, None] + off[None, :]).reshape(-1)) * dx
    ys = (((np.arange(ny) - (ny - 1) / 2)[:, None] + off[None, :]).reshape(-1)) * dy
    X, Y = xs[:, None], ys[None, :]
    img = np.zeros((nx * supersample, ny * supersample), np.float64)
    for cx, cy, a, b, rot, mu in params:
        c, s = np.cos(rot), np.sin(rot)
        u = (X - cx) * c + (Y - cy) * s
        v = -(X - cx) * s + (Y - cy) * c
        img += mu * ((u / a) ** 2 + (v / b) ** 2 <= 1.0)
    return img.reshape(nx, supersample, ny, supersample).mean(axis=(1, 3)).astype(np.float32)


def phantom(seed, nx, ny, dx, dy, n_ellipses=10, supersample=4, device=None):
    return rasterize(ellipse_params(seed, n_ellipses), nx, ny, dx, dy, supersample, device)


def phantom_batch(seeds, nx, ny, dx, dy, n_ellipses=10, supersample=4, device=None):
    if device is not None:
        import torch

        return torch.stack([phantom(s, nx, ny, dx, dy, n_ellipses, supersample, device) for s in seeds])
    return np.stack([phantom(s, nx, ny, dx, dy, n_ellipses, supersample) for s in seeds])


def ellipsoid_params(seed, n_ellipsoids=12, r_max=110.0, z_max=50.0, axis_range=(5.0, 80.0),
                     axis_z_range=(5.0, 40.0), mu_max=0.04):
    """Seeded random ellipsoids: per ellipsoid a centre at radius r_max sqrt(U),
    angle 2 pi U, height U[-z_max, z_max]; semi-axes U[5, 80] mm in-plane and
    U[5, 40] mm along z; rotation about z U[0, pi); attenuation U[0, 0.04] /mm.
    Rows (cx, cy, cz, a, b, c, rot, mu)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n_ellipsoids, 8), np.float64)
    for k in range(n_ellipsoids):
        r = r_max * np.sqrt(rng.uniform())
        phi = 2 * np.pi * rng.uniform()
        cz = rng.uniform(-z_max, z_max)
        a, b = rng.uniform(*axis_range), rng.uniform(*axis_range)
        c = rng.uniform(*axis_z_range)
        rot = rng.uniform(0, np.pi)
        mu = rng.uniform(0, mu_max)
        out[k] = [r * np.cos(phi), r * np.sin(phi), cz, a, b, c, rot, mu]
    return out


def phantom3d(seed, nx, ny, nz, dx, dy, dz, n_ellipsoids=12, supersample=2, device=None):
    """Volume [nx, ny, nz] (float32, voxel centres at (i - (n-1)/2) d) of
    additive seeded ellipsoids, supersample^3 points per voxel; torch on
    `device` if given, else numpy."""
    import torch

    dev = device if device is not None else "cpu"
    dt = torch.float32
    off = (torch.arange(supersample, dtype=dt, device=dev) + 0.5) / supersample - 0.5

    def axis(n, d):
        return ((torch.arange(n, dtype=dt, device=dev) - (n - 1) / 2)[:, None] + off[None, :]).reshape(-1) * d

    xs, ys, zs = axis(nx, dx), axis(ny, dy), axis(nz, dz)
    vol = torch.zeros((nx * supersample, ny * supersample, nz * supersample), dtype=dt, device=dev)
    X, Y = xs[:, None], ys[None, :]
    for cx, cy, cz, a, b, c, rot, mu in ellipsoid_params(seed, n_ellipsoids):
        co, si = np.cos(rot), np.sin(rot)
        u = ((X - cx) * co + (Y - cy) * si) / a
        v = (-(X - cx) * si + (Y - cy) * co) / b
        r2 = (u * u + v * v)[:, :, None] + (((zs - cz) / c) ** 2)[None, None, :]
        vol += float(mu) * (r2 <= 1.0).to(dt)
    vol = vol.reshape(nx, supersample, ny, supersample, nz, supersample).mean(dim=(1, 3, 5))
    return vol if device is not None else vol.numpy()


def add_poisson_noise(sino, i0, seed):
    """Transmission noise: N ~ Poisson(I0 exp(-p)), p~ = -ln(max(N, 1) / I0).

    Returns (noisy line integrals, counts / I0) as float32; the second array is
    the statistical weight of the data term.  Works on numpy or torch (CUDA)
    arrays; the torch path draws with a seeded generator on the array's device.
    """
    try:
        import torch
        is_torch = isinstance(sino, torch.Tensor)
    except ImportError:
        is_torch = False
    if is_torch:
        g = torch.Generator(device=sino.device)
        g.manual_seed(int(seed))
        lam = i0 * torch.exp(-sino.double())
        counts = torch.poisson(lam, generator=g)
        noisy = -torch.log(torch.clamp(counts, min=1.0) / i0)
        return noisy.float(), (counts / i0).float()
    rng = np.random.default_rng(seed)
    counts = rng.poisson(i0 * np.exp(-np.asarray(sino, np.float64)))
    noisy = -np.log(np.maximum(counts, 1) / i0)
    return noisy.astype(np.float32), (counts / i0).astype(np.float32)
