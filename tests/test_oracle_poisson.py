"""Pins the oracle's curl, calRHS and sor3D restatements (CPU only).

  * curl (kernel.cu:8352-8379): a solid-body rotation a = w(-y, x, 0) has
    |curl a| = 2w everywhere (central and one-sided differences of a linear
    field are exact); a = (0, 0, x^2) gives 2x inside and the first-order
    one-sided values dx and (2X-3)dx on the two x walls; a gradient field has
    zero curl; the mask multiplies.
  * calRHS (calRHS.cpp:7-309) with the sor3D stencil: the discrete Neumann
    problem sum_nb (p_nb - p) + rhs = 0 holds exactly for any pressure whose
    second derivatives along the axes vanish (p = xy + yz + zx + linear),
    including every face, edge and corner voxel; for p = x^2 the residual is 0
    inside and on the x faces and 2 dx^2 on the y and z faces, where the
    reference drops the tangential terms.
  * sor3D (sor3D.cpp:9-1589, damped Jacobi, weight 0.8): the in-box
    neighbour Laplacian has the eigenvectors prod_a cos(pi m_a (i_a + 1/2)/N_a)
    with eigenvalue -sum_a (2 - 2 cos(pi m_a / N_a)), so rhs = -lambda * mode
    makes the mode the fixed point, which the iteration must reach; a
    boundary band excluded by BCScheme is left untouched.
  * red-black SOR converges to the same fixed point up to a constant.
"""
import numpy as np
import pytest

import oracle as O
from omni3d import fields as F


def coords(shape, delta):
    Z, Y, X = shape
    z, y, x = np.meshgrid(np.arange(Z) * delta[2], np.arange(Y) * delta[1],
                          np.arange(X) * delta[0], indexing="ij")
    return x, y, z


def test_curl_solid_body_rotation():
    shape, delta, w = (7, 8, 9), (0.5, 0.25, 0.75), 1.5
    x, y, z = coords(shape, delta)
    c = O.curl(shape, delta, [-w * y, w * x, np.zeros(shape)])
    np.testing.assert_allclose(c, 2 * w, rtol=1e-13)


def test_curl_quadratic_component_walls():
    shape, delta = (5, 6, 9), (0.5, 1.0, 1.0)
    x, y, z = coords(shape, delta)
    c = O.curl(shape, delta, [np.zeros(shape), np.zeros(shape), x * x])
    X = shape[2]
    np.testing.assert_allclose(c[:, :, 1:-1], 2 * x[:, :, 1:-1], rtol=1e-13)
    np.testing.assert_allclose(c[:, :, 0], delta[0], rtol=1e-13)
    np.testing.assert_allclose(c[:, :, -1], (2 * X - 3) * delta[0], rtol=1e-13)


def test_curl_of_gradient_is_zero_and_mask_multiplies():
    shape = (6, 7, 8)
    f = F.quadratic(shape, seed=2, delta=(0.3, 0.4, 0.5))
    c = O.curl(shape, f["delta"], f["acc"])
    assert np.abs(c).max() < 1e-12 * max(1.0, max(np.abs(a).max() for a in f["acc"]))
    x, y, z = coords(shape, (1, 1, 1))
    rot = [-y, x, np.zeros(shape)]
    m = (np.random.default_rng(0).uniform(size=shape) > 0.3).astype(float)
    np.testing.assert_allclose(O.curl(shape, (1, 1, 1), rot, m), 2 * m, rtol=1e-13)


def inbox_residual(p, rhs):
    """sum over in-box face neighbours of (p_nb - p) + rhs, every voxel."""
    r = rhs.copy()
    for ax in range(3):
        d = np.diff(p, axis=ax)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        r[tuple(lo)] += d
        r[tuple(hi)] -= d
    return r


@pytest.mark.parametrize("shape", [(5, 6, 7), (9, 4, 6)])
def test_calrhs_exact_for_mixed_quadratic(shape):
    delta, rho = (0.5, 0.7, 0.9), 1.3
    f = F.quadratic(shape, coef=[0.3, 1.1, -0.7, 0.4, 0.0, 0.0, 0.0, 0.8, -1.2, 0.5],
                    delta=delta, rho=rho)
    rhs = O.calrhs(shape, delta, rho, f["acc"])
    r = inbox_residual(f["p"], rhs)
    assert np.abs(r).max() < 1e-12 * np.abs(rhs).max()


def test_calrhs_drops_tangential_terms_on_faces():
    shape, delta = (6, 7, 8), (0.5, 0.5, 0.5)
    f = F.quadratic(shape, coef=[0, 0, 0, 0, 1.0, 0, 0, 0, 0, 0], delta=delta)
    rhs = O.calrhs(shape, delta, 1.0, f["acc"])
    r = inbox_residual(f["p"], rhs)
    Z, Y, X = shape
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    on_x = (i == 0) | (i == X - 1)
    on_yz = (j == 0) | (j == Y - 1) | (k == 0) | (k == Z - 1)
    inner = ~(on_x | on_yz)
    np.testing.assert_allclose(r[inner], 0.0, atol=1e-12)
    np.testing.assert_allclose(r[on_x & ~on_yz], 0.0, atol=1e-12)
    # y/z-face voxels that are interior in x keep their x neighbours in the
    # stencil but lose the x term of the RHS: residual p_xx dx^2 = 2 dx^2
    sel = on_yz & ~on_x
    np.testing.assert_allclose(r[sel], 2 * delta[0] ** 2, rtol=1e-12)


def eigenmode(shape, m):
    Z, Y, X = shape
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    v = (np.cos(np.pi * m[0] * (i + 0.5) / X) * np.cos(np.pi * m[1] * (j + 0.5) / Y)
         * np.cos(np.pi * m[2] * (k + 0.5) / Z))
    lam = -sum(2 - 2 * np.cos(np.pi * m[a] / n) for a, n in enumerate((X, Y, Z)))
    return v, lam


def test_eigenmode_is_inbox_laplacian_eigenvector():
    shape = (6, 7, 8)
    v, lam = eigenmode(shape, (1, 2, 1))
    np.testing.assert_allclose(inbox_residual(v, np.zeros(shape)), lam * v, atol=1e-13)


def test_sor3d_converges_to_eigenmode():
    shape = (6, 7, 8)
    v, lam = eigenmode(shape, (1, 2, 1))
    rhs = -lam * v
    p, sweeps, residue = O.sor3d(np.zeros(shape), rhs, noitr=3000, eps=0.0)
    assert sweeps == 3000
    assert np.abs(p - v).max() < 1e-9
    # the stop rule: mean |interior residual| below eps ends the loop early
    p2, sweeps2, residue2 = O.sor3d(np.zeros(shape), rhs, noitr=3000, eps=1e-4)
    assert sweeps2 < 3000 and residue2 < 1e-4
    assert O.residual(p2, rhs) < 1e-4


def test_sor3d_bcscheme_band_untouched():
    shape = (6, 7, 8)
    rng = np.random.default_rng(4)
    p0 = rng.normal(size=shape)
    rhs = rng.normal(size=shape)
    p, sweeps, _ = O.sor3d(p0, rhs, noitr=6, eps=0.0, bcscheme=(1, 1, 1, 1, 1, 1))
    inner = (slice(1, -1),) * 3
    band = np.ones(shape, bool)
    band[inner] = False
    np.testing.assert_array_equal(p[band], p0[band])
    assert np.abs(p[inner] - p0[inner]).max() > 0


def test_redblack_reaches_the_same_fixed_point():
    shape = (6, 7, 8)
    v, lam = eigenmode(shape, (1, 2, 1))
    p = O.sor_redblack(np.zeros(shape), -lam * v, iters=600, omega=1.5)
    # Neumann problem: the solution is fixed up to a constant, which
    # Gauss-Seidel sweeps (unlike Jacobi) do not keep at zero
    d = p - v
    assert np.abs(d - d.mean()).max() < 1e-9
