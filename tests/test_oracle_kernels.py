"""Pins of the oracle's kernels (oracle/kifmm_oracle.c kblock, after
fmm3d/kernel3d.cpp:116-337) against what the mathematics fixes: the PDE each
fundamental solution satisfies, flux / jump identities, closed-form limits,
homogeneity degrees (kernel3d.cpp:104-113) and symmetry.  None of these
retypes the oracle's formulas."""
import numpy as np
import pytest

import oracle as O

MU, NU = 1.3, 0.3


def K(kt, x, n=None, coefs=(MU, NU)):
    """Kernel block at target x from a source at the origin with normal n."""
    x = np.atleast_2d(np.asarray(x, dtype=float))
    src = np.zeros((1, 3))
    nor = np.asarray(n, dtype=float).reshape(1, 3) if n is not None else src
    return O.kernel_matrix(kt, list(coefs), src, nor, x)


def fd_laplacian(f, x, h=1e-3):
    """Second-order central difference Laplacian of a vector/matrix-valued f."""
    x = np.asarray(x, float)
    out = -6.0 * f(x)
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        out = out + f(x + e) + f(x - e)
    return out / (h * h)


def fd_grad(f, x, h=1e-4):
    x = np.asarray(x, float)
    g = []
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        g.append((f(x + e) - f(x - e)) / (2 * h))
    return np.array(g)


def sphere_quadrature(n=20000):
    """Fibonacci points on the unit sphere with equal weights 4*pi/n."""
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * k
    p = np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)
    return p, np.full(n, 4 * np.pi / n)


X0 = np.array([0.31, -0.22, 0.47])


def test_dofs_and_mm_map():
    assert O.kernel_dofs(111) == (1, 1)
    assert O.kernel_dofs(301) == (4, 3)
    assert O.kernel_dofs(321) == (3, 3)
    assert O.kernel_dofs(511) == (3, 3)
    assert O.kernel_mm(111) == 111 and O.kernel_mm(321) == 301 and O.kernel_mm(511) == 511
    with pytest.raises(ValueError):
        O.kernel_dofs(999)


def test_laplace_harmonic_and_flux():
    g = lambda x: K(111, x)[0, 0]
    assert abs(fd_laplacian(g, X0)) < 1e-5 * abs(g(X0)) / np.dot(X0, X0)
    # Gauss: flux of -grad G through a sphere around the source is 1
    p, w = sphere_quadrature(4000)
    r = 0.7
    flux = 0.0
    for q, wq in zip(p, w):
        gr = fd_grad(g, r * q)
        flux += -np.dot(gr, q) * wq * r * r
    assert abs(flux - 1.0) < 1e-4


@pytest.mark.parametrize("kt", [111, 301, 321, 511])
def test_self_interaction_is_zero(kt):
    m = K(kt, [0.0, 0.0, 0.0], n=[0, 0, 1])
    assert np.all(m == 0.0)
    m = K(kt, [5e-9, 0.0, 0.0], n=[0, 0, 1])  # r < 1e-8 (kernel3d.cpp:21)
    assert np.all(m == 0.0)
    assert np.any(K(kt, [2e-8, 0.0, 0.0], n=[1, 0, 0]) != 0.0)


def test_stokeslet_solves_stokes():
    """Columns 0..2 of 301 are the Stokeslet u = G f with pressure p = f.x / (4 pi r^3):
    mu lap u = grad p and div u = 0."""
    u = lambda x: K(301, x)[:, :3]
    p = lambda x: (x / (4 * np.pi * np.linalg.norm(x) ** 3))  # p for unit forces along each axis
    lap = fd_laplacian(u, X0)
    gp = fd_grad(p, X0)  # gp[d, j] = d/dx_d p_j
    assert np.allclose(MU * lap, gp.T, atol=1e-5 * np.abs(gp).max())
    div = np.trace(fd_grad(lambda x: u(x), X0), axis1=0, axis2=1)
    assert np.all(np.abs(div) < 1e-7)
    # column 3: the potential source x / (4 pi r^3) = -grad 1/(4 pi r), harmonic and div-free
    c3 = lambda x: K(301, x)[:, 3]
    assert np.allclose(c3(X0), -fd_grad(lambda x: K(111, x)[0, 0], X0), rtol=1e-6)
    assert abs(np.trace(fd_grad(c3, X0))) < 1e-6


def test_stokes_dl_jump_identity():
    """Double-layer identity: the integral of the 321 kernel over the unit sphere
    with outward normals is the identity inside and zero outside."""
    p, w = sphere_quadrature(20000)
    for x, expect in [(np.array([0.1, -0.2, 0.15]), np.eye(3)), (np.array([1.6, 0.3, -0.4]), np.zeros((3, 3)))]:
        m = O.kernel_matrix(321, [1.0], p, p, x[None, :])  # 3 x (3 n)
        tot = sum(m[:, 3 * j:3 * j + 3] * w[j] for j in range(len(w)))
        assert np.allclose(tot, expect, atol=2e-3), tot


def test_stokes_dl_divergence_free():
    n = np.array([0.3, -0.5, 0.81])
    n = n / np.linalg.norm(n)
    f = lambda x: K(321, x, n)
    div = np.einsum("ddj->j", fd_grad(f, X0))
    assert np.all(np.abs(div) < 1e-6 * np.abs(f(X0)).max())


def test_navier_solves_lame():
    """Kelvin solution: mu lap u + (lambda + mu) grad div u = 0, lambda = 2 mu nu / (1 - 2 nu)."""
    lam = 2 * MU * NU / (1 - 2 * NU)
    u = lambda x: K(511, x)
    lap = fd_laplacian(u, X0, h=2e-3)
    # grad div u: div over the first index of u (target components)
    div = lambda x: np.einsum("ddj->j", fd_grad(u, x, h=1e-4))
    gd = fd_grad(div, X0, h=1e-3)  # gd[d, j]
    res = MU * lap + (lam + MU) * gd
    assert np.abs(res).max() < 2e-4 * np.abs(MU * lap).max()


def test_navier_incompressible_limit_is_stokeslet():
    m = O.kernel_matrix(511, [MU, 0.5], np.zeros((1, 3)), np.zeros((1, 3)), X0[None, :])
    s = O.kernel_matrix(301, [MU], np.zeros((1, 3)), np.zeros((1, 3)), X0[None, :])[:, :3]
    assert np.allclose(m, s, rtol=1e-14)


@pytest.mark.parametrize("kt,deg", [(111, [1]), (301, [1, 1, 1, 2]), (511, [1, 1, 1])])
def test_homogeneity_degrees(kt, deg):
    a = 2.0 ** -3
    k1 = K(kt, X0)
    k2 = K(kt, a * X0)
    for s, d in enumerate(deg):
        assert np.allclose(k2[:, s], a ** -d * k1[:, s], rtol=1e-14)


@pytest.mark.parametrize("kt", [111, 511])
def test_reciprocity(kt):
    assert np.allclose(K(kt, X0), K(kt, -X0).T, rtol=1e-15)
