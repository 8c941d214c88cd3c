"""Seeded synthetic phantoms and noise for the benchmark workloads.

Attenuation is in 1/mm.  Volumes are returned in the C-order layout
vol[y, x, z] (== the reference's vol[(y*XN + x)*SLN + z]); sinograms as
prj[v, u, z].
"""
from __future__ import annotations

import numpy as np

# Modified Shepp-Logan (Toft): value, semi-axes a, b, centre x, y, angle (deg)
_SHEPP = [
    (1.0, 0.69, 0.92, 0.0, 0.0, 0.0),
    (-0.8, 0.6624, 0.8740, 0.0, -0.0184, 0.0),
    (-0.2, 0.11, 0.31, 0.22, 0.0, -18.0),
    (-0.2, 0.16, 0.41, -0.22, 0.0, 18.0),
    (0.1, 0.21, 0.25, 0.0, 0.35, 0.0),
    (0.1, 0.046, 0.046, 0.0, 0.1, 0.0),
    (0.1, 0.046, 0.046, 0.0, -0.1, 0.0),
    (0.1, 0.046, 0.023, -0.08, -0.605, 0.0),
    (0.1, 0.023, 0.023, 0.0, -0.606, 0.0),
    (0.1, 0.023, 0.046, 0.06, -0.605, 0.0),
]


def _grid(XN, YN, dx, ss=2):
    """Supersampled pixel-centre coordinates (mm), shape (YN, XN, ss*ss)."""
    o = (np.arange(ss) + 0.5) / ss - 0.5
    xs = ((np.arange(XN) - (XN - 1) / 2.0)[:, None] + o[None, :]) * dx  # (XN, ss)
    ys = ((np.arange(YN) - (YN - 1) / 2.0)[:, None] + o[None, :]) * dx
    X = np.broadcast_to(xs[None, :, None, :], (YN, XN, ss, ss)).reshape(YN, XN, ss * ss)
    Y = np.broadcast_to(ys[:, None, :, None], (YN, XN, ss, ss)).reshape(YN, XN, ss * ss)
    return X, Y


def _ellipse_mask(X, Y, cx, cy, a, b, ang):
    c, s = np.cos(ang), np.sin(ang)
    xr = (X - cx) * c + (Y - cy) * s
    yr = -(X - cx) * s + (Y - cy) * c
    return (xr / a) ** 2 + (yr / b) ** 2 <= 1.0


def shepp_logan_2d(XN, YN, dx, fov_radius, mu_water=0.0195, n_random=20, seed=0,
                   mu_range=(0.01, 0.03), ss=2):
    """Shepp-Logan scaled to water = mu_water plus ``n_random`` seeded random
    ellipses of attenuation U(mu_range), replacing the values they cover."""
    X, Y = _grid(XN, YN, dx, ss)
    img = np.zeros(X.shape, np.float64)
    R = fov_radius
    for v, a, b, cx, cy, ang in _SHEPP:
        img += v * mu_water * _ellipse_mask(X, Y, cx * R, cy * R, a * R, b * R, np.deg2rad(ang))
    rng = np.random.default_rng(seed)
    for _ in range(n_random):
        r = rng.uniform(0.0, 0.6) * R
        t = rng.uniform(0, 2 * np.pi)
        a = rng.uniform(0.02, 0.12) * R
        b = rng.uniform(0.02, 0.12) * R
        m = _ellipse_mask(X, Y, r * np.cos(t), r * np.sin(t), a, b, rng.uniform(0, np.pi))
        img[m] = rng.uniform(*mu_range)
    return img.mean(axis=2).astype(np.float32)


def ellipsoid_phantom(XN, YN, SLN, dx, dz=None, n=24, seed=1, mu_range=(0.0, 0.03),
                      fov_mask=True):
    """3-D seeded random-ellipsoid phantom vol[y, x, z] inside a circular FOV:
    an elliptical body of water (0.019/mm) holding ``n`` random ellipsoids of
    attenuation U(mu_range).  Pixel-centre sampling, bounding-box limited."""
    dz = dx if dz is None else dz
    rng = np.random.default_rng(seed)
    xs = (np.arange(XN) - (XN - 1) / 2.0) * dx
    ys = (np.arange(YN) - (YN - 1) / 2.0) * dx
    zc = (np.arange(SLN) - (SLN - 1) / 2.0) * dz
    R = 0.5 * min(XN, YN) * dx
    H = 0.5 * SLN * dz
    X, Y = np.meshgrid(xs, ys)
    body = (((X / (0.85 * R)) ** 2 + (Y / (0.7 * R)) ** 2) <= 1.0).astype(np.float32) * 0.019
    vol = np.repeat(body[:, :, None], SLN, axis=2)
    for _ in range(n):
        cx, cy = rng.uniform(-0.5, 0.5, 2) * R
        cz = rng.uniform(-1, 1) * H
        a, b = rng.uniform(0.05, 0.25, 2) * R
        c = rng.uniform(0.2, 1.0) * max(H, dz)
        ang = rng.uniform(0, np.pi)
        mu = np.float32(rng.uniform(*mu_range))
        r = max(a, b)
        x0, x1 = np.searchsorted(xs, cx - r), np.searchsorted(xs, cx + r)
        y0, y1 = np.searchsorted(ys, cy - r), np.searchsorted(ys, cy + r)
        if x1 <= x0 or y1 <= y0:
            continue
        Xs, Ys = X[y0:y1, x0:x1], Y[y0:y1, x0:x1]
        co, si = np.cos(ang), np.sin(ang)
        xr = (Xs - cx) * co + (Ys - cy) * si
        yr = -(Xs - cx) * si + (Ys - cy) * co
        q = (xr / a) ** 2 + (yr / b) ** 2
        for z in range(SLN):
            w = 1.0 - ((zc[z] - cz) / c) ** 2
            if w > 0:
                sub = vol[y0:y1, x0:x1, z]
                sub[q <= w] = mu
    if fov_mask:
        vol *= (np.sqrt(X ** 2 + Y ** 2) <= 0.95 * R)[:, :, None]
    return np.ascontiguousarray(vol, np.float32)


def poisson_sinogram(prj, dx, I0, seed=2):
    """Noisy sinogram in the projector's units (line integral / dx):
    counts ~ Poisson(I0 exp(-prj*dx)), p = -ln(max(counts, 1)/I0)/dx."""
    rng = np.random.default_rng(seed)
    lam = I0 * np.exp(-np.asarray(prj, np.float64) * dx)
    c = rng.poisson(lam)
    return (-np.log(np.maximum(c, 1) / I0) / dx).astype(np.float32)


def uniform_angles(PN, SLN, start=0.0, span=2 * np.pi):
    """hangs[v, z] = start + span*v/PN for every slice (float32)."""
    a = (start + span * np.arange(PN) / PN).astype(np.float32)
    return np.ascontiguousarray(np.repeat(a[:, None], SLN, axis=1))



def gaussian_count_sinogram(prj, dx, I0, seed=2, threads=None):
    """Fast large-array stand-in for :func:`poisson_sinogram`: counts
    lam + sqrt(lam)*N(0,1) with lam = I0 exp(-prj*dx) (the normal
    approximation of Poisson(lam)), p = -ln(max(counts, 1)/I0)/dx, in float32.
    Views are split into chunks, each with its own child seed of ``seed``, and
    processed on a thread pool (numpy releases the GIL in these loops)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    p = np.asarray(prj, np.float32)
    out = np.empty_like(p)
    nchunk = max(1, min(64, p.shape[0]))
    bounds = np.linspace(0, p.shape[0], nchunk + 1).astype(int)
    seeds = np.random.SeedSequence(seed).spawn(nchunk)
    f_dx, f_I0 = np.float32(dx), np.float32(I0)

    def work(i):
        a, b = bounds[i], bounds[i + 1]
        rng = np.random.default_rng(seeds[i])
        lam = f_I0 * np.exp(-p[a:b] * f_dx)
        c = lam + np.sqrt(lam) * rng.standard_normal(lam.shape, dtype=np.float32)
        np.maximum(c, np.float32(1.0), out=c)
        out[a:b] = -np.log(c / f_I0) / f_dx

    n = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(work, range(nchunk)))
    return out


# ------------------------------------------------------------------ workloads
MIX_XN = (256, 384, 512, 640, 768)
MIX_PN = (720, 1152, 1536, 2304)


def mix_specs(n, seed):
    """Sizes of the uneven reconstruction mix: image 256^2-768^2 over a 500 mm
    field of view, 8-128 slices, 720-2304 views, I0 log-uniform in
    [5e4, 5e5]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        XN = int(rng.choice(MIX_XN))
        out.append({"XN": XN, "SLN": int(rng.integers(8, 129)), "PN": int(rng.choice(MIX_PN)),
                    "I0": float(np.exp(rng.uniform(np.log(5e4), np.log(5e5))))})
    return out


def helical_cylinder_counts(W, H, DefTimes, rotations, SO, step, detc, I0=1e5, gain=0.5,
                            radius=200.0, mu=0.019, seed=5, noise=True, threads=None):
    """Raw helical data raw[view, ch, row] (uint16) of a water cylinder of
    `radius` mm on the rotation axis: expected counts gain*I0*exp(-mu*chord)
    per channel (arc detector of `step`-radian channels centred on 1-based
    channel `detc`, source radius SO), modulated by 5% within each rotation
    and by a seeded +-2% row gain; with `noise`, every sample gets the normal
    approximation of Poisson noise (lam + sqrt(lam)*N(0,1), each rotation
    from its own child seed); rounded and clipped to [1, 65535] (the gain
    keeps I0 = 1e5 inside 16 bits)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    beta = (np.arange(W) - (detc - 1)) * step
    d = SO * np.abs(np.sin(beta))
    chord = 2.0 * np.sqrt(np.maximum(radius ** 2 - d ** 2, 0.0))
    base = gain * I0 * np.exp(-mu * chord)
    rng = np.random.default_rng(seed)
    mod = 1.0 + 0.05 * np.sin(6 * np.pi * np.arange(DefTimes) / DefTimes)
    lam = (base[None, :, None] * mod[:, None, None] *
           rng.uniform(0.98, 1.02, (1, 1, H))).astype(np.float32)
    TV = DefTimes * rotations
    raw = np.empty((TV, W, H), np.uint16)
    if not noise:
        one = np.clip(np.rint(lam), 1, 65535).astype(np.uint16)
        for r0 in range(0, TV, DefTimes):
            raw[r0:r0 + DefTimes] = one
        return raw
    seeds = np.random.SeedSequence(seed).spawn(rotations)
    sd = np.sqrt(lam)

    def work(r):
        g = np.random.default_rng(seeds[r])
        c = lam + sd * g.standard_normal(lam.shape, dtype=np.float32)
        raw[r * DefTimes:(r + 1) * DefTimes] = np.clip(np.rint(c), 1, 65535)

    n = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(work, range(rotations)))
    return raw
