"""GPU checks of the adjoint pair omni3d_forward / omni3d_backward (the
linear map g = grad p -> p of the plain and plain-mask methods from a zero
pressure, K iterations):

  * transpose identity <y, F g> = <F^T y, g> in FP64 (rel <= 1e-12);
  * against the explicit FP64 Jacobian of the oracle on a tiny grid (every
    column is one oracle reconstruction of a unit gradient), rel <= 1e-12;
  * central finite differences on 64 random unit directions in grad-p space
    at the masked Fourier-sphere workload (32^3), rel <= 1e-6;
  * the source-restricted backward (Horner steps over the source voxels --
    ray entries for plain, run starts for plain-mask -- with the transposed
    source rows) against the full-volume Horner sum, rel <= 1e-12, at 48^3
    with 162 directions;
  * FP32-Kahan forward/backward within 1e-5 of FP64;
  * a torch.autograd.Function over the pair fits a gradient correction to
    512 sparse pressure probes with Tikhonov lambda = 1e-3 (50 steps).
"""
import numpy as np
import pytest

import oracle as O
import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


class Adjoint:
    def __init__(self, shape, dirs, delta, method, K, mask=None, precision=omni3d.FP64,
                 solver=omni3d.NAIVE):
        import torch
        self.shape, self.V = shape, int(np.prod(shape))
        self.dt = torch.float64 if precision == omni3d.FP64 else torch.float32
        self.ctx = omni3d.Context(shape, delta, method=method, max_iterations=K,
                                  precision=precision, solver=solver)
        self.ctx.set_directions(dirs)
        z = [np.zeros(shape)] * 3
        self.ctx.set_field(z, mask=mask)

    def forward(self, g):
        import torch
        g = g.to(self.dt).contiguous()
        p = torch.empty(self.V, dtype=self.dt, device="cuda")
        self.ctx.forward_device(g.data_ptr(), p.data_ptr())
        torch.cuda.synchronize()
        return p

    def backward(self, y):
        import torch
        y = y.to(self.dt).contiguous()
        dg = torch.empty(3 * self.V, dtype=self.dt, device="cuda")
        self.ctx.backward_device(y.data_ptr(), dg.data_ptr())
        torch.cuda.synchronize()
        return dg


def sphere(shape, r):
    Z, Y, X = shape
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    c = [(X - 1) / 2, (Y - 1) / 2, (Z - 1) / 2]
    return (np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) > r).astype(np.float64)


SOLVERS = [(omni3d.PLAIN, omni3d.NAIVE), (omni3d.PLAIN, omni3d.AFFINE),
           (omni3d.PLAIN_MASK, omni3d.NAIVE), (omni3d.PLAIN_MASK, omni3d.AFFINE)]


@pytest.mark.parametrize("method,solver", SOLVERS)
def test_transpose_identity(gpu_available, ref_dirs2, method, solver):
    import torch
    shape = (16, 17, 18)
    print(f"case adjoint transpose method={method} solver={solver}", flush=True)
    m = sphere(shape, 4.0) if method == omni3d.PLAIN_MASK else None
    A = Adjoint(shape, ref_dirs2, (0.3, 0.4, 0.5), method, 6, mask=m, solver=solver)
    gen = torch.Generator(device="cuda").manual_seed(0)
    g = torch.randn(3 * A.V, dtype=torch.float64, device="cuda", generator=gen)
    y = torch.randn(A.V, dtype=torch.float64, device="cuda", generator=gen)
    lhs = float(torch.dot(y, A.forward(g)))
    rhs = float(torch.dot(A.backward(y), g))
    print("transpose identity", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1e-300) * 10


@pytest.mark.parametrize("method,solver", SOLVERS)
def test_against_oracle_jacobian(gpu_available, ref_dirs2, method, solver):
    import torch
    shape = (5, 6, 7)
    print(f"case adjoint vs oracle jacobian method={method} solver={solver}", flush=True)
    V = int(np.prod(shape))
    delta, K = (0.5, 0.6, 0.7), 4
    m = None
    if method == omni3d.PLAIN_MASK:
        m = np.ones(shape)
        m[2, 2:4, 2:5] = 0.0
    dirs = ref_dirs2[::3]
    J = np.zeros((V, 3 * V))
    for col in range(3 * V):
        g = np.zeros(3 * V)
        g[col] = 1.0
        # the oracle integrates -rho a with rho = 1: a = -g
        acc = [-g[c * V:(c + 1) * V].reshape(shape) for c in range(3)]
        p, _, _, _ = O.reconstruct(shape, dirs, 1.0, delta, 1.0, acc, curl=m,
                                   method=O.PLAIN if m is None else O.PLAIN_MASK,
                                   max_iter=K, seed_pn=0)
        J[:, col] = p.reshape(-1)
    A = Adjoint(shape, dirs, delta, method, K, mask=m, solver=solver)
    rng = np.random.default_rng(3)
    y = rng.normal(size=V)
    got = A.backward(torch.from_numpy(y).cuda()).cpu().numpy()
    want = J.T @ y
    assert F.rel_l2(got, want) < 1e-12
    g = rng.normal(size=3 * V)
    pf = A.forward(torch.from_numpy(g).cuda()).cpu().numpy()
    assert F.rel_l2(pf, J @ g) < 1e-12


def test_finite_differences_fourier_sphere(gpu_available, ref_dirs3):
    import torch
    n = 32
    print("case adjoint finite differences", flush=True)
    f = F.fourier_sphere(n, seed=1)
    shape = (n, n, n)
    A = Adjoint(shape, ref_dirs3, f["delta"], omni3d.PLAIN_MASK, 5, mask=f["mask"])
    g0 = torch.from_numpy(np.concatenate([c.reshape(-1) for c in f["grad"]])).cuda()
    gen = torch.Generator(device="cuda").manual_seed(1)
    y = torch.randn(A.V, dtype=torch.float64, device="cuda", generator=gen)
    dLdg = A.backward(y)
    # F is linear, so a central difference is exact up to rounding for any
    # step; the step is 1 % of |g0| so that the difference of two large
    # reconstructions does not cancel to the rounding level
    eps = 1e-2 * float(torch.linalg.norm(g0))
    worst = 0.0
    for i in range(64):
        d = torch.randn(3 * A.V, dtype=torch.float64, device="cuda", generator=gen)
        d /= torch.linalg.norm(d)
        fd = float(torch.dot(y, A.forward(g0 + eps * d) - A.forward(g0 - eps * d))) / (2 * eps)
        an = float(torch.dot(dLdg, d))
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    print("finite-difference worst rel error", worst)
    assert worst < 1e-6


@pytest.mark.parametrize("method,K", [(omni3d.PLAIN, 1), (omni3d.PLAIN, 2), (omni3d.PLAIN, 3),
                                      (omni3d.PLAIN, 25), (omni3d.PLAIN_MASK, 2),
                                      (omni3d.PLAIN_MASK, 25)])
def test_source_restricted_backward(gpu_available, ref_dirs3, method, K):
    import torch
    n = 48
    print(f"case adjoint source-restricted vs full-volume Horner method={method} K={K}",
          flush=True)
    f = F.fourier_sphere(n, seed=3)
    shape = (n, n, n)
    m = f["mask"] if method == omni3d.PLAIN_MASK else None
    full = Adjoint(shape, ref_dirs3, f["delta"], method, K, mask=m)
    src = Adjoint(shape, ref_dirs3, f["delta"], method, K, mask=m, solver=omni3d.AFFINE)
    gen = torch.Generator(device="cuda").manual_seed(5)
    y = torch.randn(full.V, dtype=torch.float64, device="cuda", generator=gen)
    a, b = src.backward(y), full.backward(y)
    err = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    print("source-restricted vs full-volume rel-L2", err)
    assert err < 1e-12
    # a second call reuses the transposed rows kept with the plan (the final
    # scatter adds in FP64 atomics, so the order of its last bits may vary)
    again = float(torch.linalg.norm(src.backward(y) - a) / torch.linalg.norm(a))
    assert again < 1e-13


def test_masked_backward_cache_follows_the_mask(gpu_available, ref_dirs3):
    """The plain-mask backward keeps its per-mask visit counts and run-start
    rows between calls: a forward in between keeps them, a new mask set with
    set_field rebuilds them (the result equals a fresh context's)."""
    import torch
    n = 40
    print("case adjoint plain-mask cache", flush=True)
    f = F.fourier_sphere(n, seed=4)
    shape = (n, n, n)
    m1 = f["mask"]
    m2 = m1.copy()
    m2[: n // 3] = 1.0
    m2[:, : n // 4] = 0.0
    gen = torch.Generator(device="cuda").manual_seed(9)
    y = torch.randn(n ** 3, dtype=torch.float64, device="cuda", generator=gen)
    g = torch.randn(3 * n ** 3, dtype=torch.float64, device="cuda", generator=gen)
    A = Adjoint(shape, ref_dirs3, f["delta"], omni3d.PLAIN_MASK, 6, mask=m1, solver=omni3d.AFFINE)
    b1 = A.backward(y)
    A.forward(g)
    b1b = A.backward(y)
    assert float(torch.linalg.norm(b1b - b1) / torch.linalg.norm(b1)) < 1e-13
    A.ctx.set_field([np.zeros(shape)] * 3, mask=m2)
    b2 = A.backward(y)
    fresh = Adjoint(shape, ref_dirs3, f["delta"], omni3d.PLAIN_MASK, 6, mask=m2,
                    solver=omni3d.AFFINE)
    ref = fresh.backward(y)
    assert float(torch.linalg.norm(b2 - ref) / torch.linalg.norm(ref)) < 1e-13
    assert float(torch.linalg.norm(b2 - b1) / torch.linalg.norm(b1)) > 1e-3
    # the mask removed: the rule's counts and rows follow (all-ones mask)
    A.ctx.set_field([np.zeros(shape)] * 3)
    ones = Adjoint(shape, ref_dirs3, f["delta"], omni3d.PLAIN_MASK, 6, mask=np.ones(shape),
                   solver=omni3d.AFFINE)
    b3, ref3 = A.backward(y), ones.backward(y)
    assert float(torch.linalg.norm(b3 - ref3) / torch.linalg.norm(ref3)) < 1e-13


def test_finite_differences_plain_affine(gpu_available, ref_dirs3):
    import torch
    n = 32
    print("case adjoint finite differences plain affine", flush=True)
    f = F.fourier_sphere(n, seed=1)
    shape = (n, n, n)
    A = Adjoint(shape, ref_dirs3, f["delta"], omni3d.PLAIN, 30, solver=omni3d.AFFINE)
    g0 = torch.from_numpy(np.concatenate([c.reshape(-1) for c in f["grad"]])).cuda()
    gen = torch.Generator(device="cuda").manual_seed(6)
    y = torch.randn(A.V, dtype=torch.float64, device="cuda", generator=gen)
    dLdg = A.backward(y)
    eps = 1e-2 * float(torch.linalg.norm(g0))
    worst = 0.0
    for i in range(64):
        d = torch.randn(3 * A.V, dtype=torch.float64, device="cuda", generator=gen)
        d /= torch.linalg.norm(d)
        fd = float(torch.dot(y, A.forward(g0 + eps * d) - A.forward(g0 - eps * d))) / (2 * eps)
        an = float(torch.dot(dLdg, d))
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    print("finite-difference worst rel error", worst)
    assert worst < 1e-6


def test_fp32_close_to_fp64(gpu_available, ref_dirs2):
    import torch
    shape = (20, 20, 20)
    print("case adjoint fp32", flush=True)
    m = sphere(shape, 5.0)
    A64 = Adjoint(shape, ref_dirs2, (0.3, 0.3, 0.3), omni3d.PLAIN_MASK, 5, mask=m)
    A32 = Adjoint(shape, ref_dirs2, (0.3, 0.3, 0.3), omni3d.PLAIN_MASK, 5, mask=m,
                  precision=omni3d.FP32_KAHAN)
    gen = torch.Generator(device="cuda").manual_seed(2)
    g = torch.randn(3 * A64.V, dtype=torch.float64, device="cuda", generator=gen)
    y = torch.randn(A64.V, dtype=torch.float64, device="cuda", generator=gen)
    p64, p32 = A64.forward(g), A32.forward(g).double()
    assert float(torch.linalg.norm(p32 - p64) / torch.linalg.norm(p64)) < 1e-5
    d64, d32 = A64.backward(y), A32.backward(y).double()
    assert float(torch.linalg.norm(d32 - d64) / torch.linalg.norm(d64)) < 1e-6


def test_probe_fit(gpu_available, ref_dirs2):
    """Fit delta(grad p) so that F(g + delta) matches 512 sparse probes of the
    true pressure, with Tikhonov lambda = 1e-3, 50 Adam steps."""
    import torch

    n = 24
    print("case probe fit", flush=True)
    f = F.fourier_sphere(n, seed=2, noise=0.05, outlier_frac=0.0)
    shape = (n, n, n)
    A = Adjoint(shape, ref_dirs2, f["delta"], omni3d.PLAIN_MASK, 8, mask=f["mask"])

    class Recon(torch.autograd.Function):
        @staticmethod
        def forward(ctx, g):
            return A.forward(g)

        @staticmethod
        def backward(ctx, dp):
            return A.backward(dp)

    g_noisy = torch.from_numpy(np.concatenate([c.reshape(-1) for c in f["grad"]])).cuda()
    p_true = torch.from_numpy(f["p"].reshape(-1)).cuda()
    rng = np.random.default_rng(4)
    valid = np.nonzero(f["mask"].reshape(-1) > 0)[0]
    probes = torch.from_numpy(rng.choice(valid, 512, replace=False)).cuda()
    delta = torch.zeros_like(g_noisy, requires_grad=True)
    opt = torch.optim.Adam([delta], lr=0.05)
    lam = 1e-3
    losses = []
    for step in range(50):
        opt.zero_grad()
        p = Recon.apply(g_noisy + delta)
        # pressure is defined up to a constant: compare mean-removed probes
        r = p[probes] - p_true[probes]
        r = r - r.mean()
        loss = (r * r).mean() + lam * (delta * delta).mean()
        loss.backward()
        opt.step()
        losses.append(float(loss))
    print("probe-fit loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.5 * losses[0]
