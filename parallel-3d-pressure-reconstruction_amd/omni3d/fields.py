"""Seeded synthetic pressure fields with closed-form gradients.

Arrays are shaped (Z, Y, X) so that the C order index is i + j*X + k*X*Y,
the layout of the reference (kernel.cu:1760).  The reconstruction input is
material acceleration a = -grad(p)/rho (kernel.cu:1816 integrates -rho*a).
Each generator returns a dict with p (exact pressure), acc (ax, ay, az),
mask, delta (dx, dy, dz) and rho.
"""
import numpy as np


def _grid(shape, h):
    Z, Y, X = shape
    z, y, x = np.meshgrid(np.arange(Z) * h[2], np.arange(Y) * h[1], np.arange(X) * h[0],
                          indexing="ij")
    return x, y, z


def _noise(rng, g, sigma_frac):
    """Per-component N(0, (s*rms)^2) with rms = sqrt(mean |grad p|^2)."""
    if sigma_frac <= 0:
        return g
    rms = np.sqrt(np.mean(g[0] ** 2 + g[1] ** 2 + g[2] ** 2))
    return [c + rng.normal(0.0, sigma_frac * rms, c.shape) for c in g]


def _pack(p, g, h, rho=1.0, mask=None):
    acc = [(-c / rho) for c in g]
    return {"p": p, "acc": acc, "grad": g, "delta": tuple(float(v) for v in h), "rho": rho,
            "mask": np.ones(p.shape) if mask is None else mask}


def taylor_green(n=32, noise=0.05, seed=0, shape=None, phase=(0.0, 0.0, 0.0), dtype=np.float64):
    """p = (1/16)(cos 2x + cos 2y)(cos 2z + 2) on x_i = i*h, h = 2 pi / n."""
    shape = shape or (n, n, n)
    h = 2 * np.pi / n
    x, y, z = _grid(shape, (h, h, h))
    x, y, z = x + phase[0], y + phase[1], z + phase[2]
    p = (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2) / 16.0
    gx = -2 * np.sin(2 * x) * (np.cos(2 * z) + 2) / 16.0
    gy = -2 * np.sin(2 * y) * (np.cos(2 * z) + 2) / 16.0
    gz = -2 * (np.cos(2 * x) + np.cos(2 * y)) * np.sin(2 * z) / 16.0
    rng = np.random.default_rng(seed)
    g = _noise(rng, [gx, gy, gz], noise)
    out = _pack(p, [c.astype(dtype) for c in g], (h, h, h))
    return out


def fourier_sphere(n=128, modes=64, kmax=8, noise=0.03, outlier_frac=0.005, seed=0,
                   sphere_radius=0.2):
    """Random Fourier sum p = sum A_k cos(k.x + phi_k), 1 <= |k| <= kmax integer
    wavevectors, A_k ~ |k|^(-5/3), on L = 2 pi; a centred sphere of radius
    0.2 L is masked out (mask 0); 3 % noise and 0.5 % outliers at 10x rms."""
    rng = np.random.default_rng(seed)
    L = 2 * np.pi
    h = L / n
    x, y, z = _grid((n, n, n), (h, h, h))
    ks = []
    while len(ks) < modes:
        k = rng.integers(-kmax, kmax + 1, 3)
        kk = np.sqrt((k * k).sum())
        if 1 <= kk <= kmax:
            ks.append(k)
    p = np.zeros((n, n, n))
    gx, gy, gz = np.zeros_like(p), np.zeros_like(p), np.zeros_like(p)
    for k in ks:
        amp = np.sqrt((k * k).sum()) ** (-5.0 / 3.0)
        ph = rng.uniform(0, 2 * np.pi)
        arg = k[0] * x + k[1] * y + k[2] * z + ph
        p += amp * np.cos(arg)
        s = -amp * np.sin(arg)
        gx += k[0] * s
        gy += k[1] * s
        gz += k[2] * s
    c = (n - 1) * h / 2
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    mask = (r > sphere_radius * L).astype(np.float64)
    g = _noise(rng, [gx, gy, gz], noise)
    rms = np.sqrt(np.mean(gx ** 2 + gy ** 2 + gz ** 2))
    nout = int(round(outlier_frac * p.size))
    idx = rng.choice(p.size, nout, replace=False)
    for comp in g:
        comp.reshape(-1)[idx] = rng.normal(0.0, 10 * rms, nout)
    return _pack(p, g, (h, h, h), mask=mask)


def burgers_vortex(n=256, noise=0.02, seed=0, rc=0.1, gamma=1.0, rho=1.0, L=1.0):
    """Swirl-only Burgers / Lamb-Oseen vortex along z through the grid centre:
    p(r) = -(rho G^2 / (8 pi^2 rc^2)) [(1 - e^-eta)^2 / eta + 2 (E1(eta) - E1(2 eta))],
    eta = r^2/rc^2, dp/dr = rho v^2 / r, v = G/(2 pi r) (1 - e^-eta)."""
    from scipy.special import exp1

    h = L / n
    x, y, _ = _grid((n, n, n), (h, h, h))
    c = (n - 1) * h / 2
    rx, ry = x - c, y - c
    r2 = rx * rx + ry * ry
    eta = np.maximum(r2 / (rc * rc), 1e-300)
    one = -np.expm1(-eta)
    p = -(rho * gamma ** 2 / (8 * np.pi ** 2 * rc ** 2)) * (
        one * one / eta + 2 * (exp1(eta) - exp1(2 * eta)))
    # dp/dr / r = rho v^2 / r^2 = rho G^2 (1 - e^-eta)^2 / (4 pi^2 r^4)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = rho * gamma ** 2 * one * one / (4 * np.pi ** 2 * r2 * r2)
    f = np.where(r2 > 0, f, 0.0)
    gx, gy, gz = f * rx, f * ry, np.zeros_like(p)
    rng = np.random.default_rng(seed)
    g = _noise(rng, [gx, gy, gz], noise)
    return _pack(p, g, (h, h, h), rho=rho)


def quadratic(shape, coef=None, seed=0, delta=(1.0, 1.0, 1.0), rho=1.0):
    """Quadratic p (linear gradient): the trapezoid rule integrates it exactly."""
    h = delta
    rng = np.random.default_rng(seed)
    c = rng.normal(size=10) if coef is None else np.asarray(coef, dtype=np.float64)
    x, y, z = _grid(shape, h)
    p = (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * x + c[5] * y * y + c[6] * z * z
         + c[7] * x * y + c[8] * y * z + c[9] * x * z)
    gx = c[1] + 2 * c[4] * x + c[7] * y + c[9] * z
    gy = c[2] + 2 * c[5] * y + c[7] * x + c[8] * z
    gz = c[3] + 2 * c[6] * z + c[8] * y + c[9] * x
    return _pack(p, [gx, gy, gz], h, rho=rho)


def snapshot_series(count=64, seed=0):
    """Grid shapes (Z, Y, X) drawn by seed from {64^3, 96^3, 128x96x64,
    160x128x96} with a Taylor-Green field, random phase and 0-5 % noise."""
    rng = np.random.default_rng(seed)
    sizes = [(64, 64, 64), (96, 96, 96), (64, 96, 128), (96, 128, 160)]
    out = []
    for i in range(count):
        Z, Y, X = sizes[rng.integers(0, 4)]
        ph = tuple(rng.uniform(0, 2 * np.pi, 3))
        nz = rng.uniform(0.0, 0.05)
        out.append({"shape": (Z, Y, X), "phase": ph, "noise": nz, "seed": int(rng.integers(1 << 30))})
    return out


def rel_l2(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        m = np.asarray(mask) != 0
        a, b = a[m], b[m]
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def demean(a, mask=None):
    a = np.asarray(a, dtype=np.float64)
    if mask is None:
        return a - a.mean()
    m = np.asarray(mask) != 0
    out = a.copy()
    out[m] -= a[m].mean()
    return out
