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
    # kernel3d.cpp:24-85 (dofs), fmm3d_setup.cpp:41-54 (translation kernels)
    dofs = {111: (1, 1), 121: (1, 1), 301: (4, 3), 311: (3, 3), 312: (3, 1), 321: (3, 3), 322: (3, 1),
            511: (3, 3), 521: (3, 3)}
    for kt, d in dofs.items():
        assert O.kernel_dofs(kt) == d
    mm = {111: 111, 121: 111, 311: 301, 312: 111, 321: 301, 322: 111, 511: 511, 521: 511}
    for kt, m in mm.items():
        assert O.kernel_mm(kt) == m
        # the user kernel's target dof equals the translation kernel's (check surfaces are shared)
        assert O.kernel_dofs(kt)[1] == O.kernel_dofs(m)[1]
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


@pytest.mark.parametrize("kt", [111, 121, 301, 311, 312, 321, 322, 511, 521])
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


# ---- the other kernel types of the reference (kernel3d.cpp:139-154, 185-221, 243-265, 338-369)

def test_laplace_double_layer_is_normal_derivative_and_jump():
    """121 = n . grad_x of the single layer (source at the origin, normal n), harmonic,
    and its integral over the unit sphere with outward normals is 1 inside, 0 outside."""
    n = np.array([0.2, -0.6, 0.77])
    n = n / np.linalg.norm(n)
    d = K(121, X0, n)[0, 0]
    g = fd_grad(lambda x: K(111, x)[0, 0], X0)
    assert abs(d - np.dot(n, g)) < 1e-7 * abs(d)
    assert abs(fd_laplacian(lambda x: K(121, x, n)[0, 0], X0)) < 1e-4 * abs(d) / np.dot(X0, X0)
    p, w = sphere_quadrature(20000)
    for x, expect in [(np.array([0.1, -0.2, 0.15]), 1.0), (np.array([1.6, 0.3, -0.4]), 0.0)]:
        m = O.kernel_matrix(121, [], p, p, x[None, :])
        assert abs(np.dot(m[0], w) - expect) < 2e-3


def test_stokes_single_layer_velocity_pressure_pair():
    """311 is the Stokeslet (the velocity block of 301) and (u, p) = (311, 312)
    solve mu lap u = grad p, div u = 0."""
    assert np.allclose(K(311, X0), K(301, X0)[:, :3], rtol=1e-15, atol=0)
    u = lambda x: K(311, x)
    p = lambda x: K(312, x)[0]  # p[j]: pressure of a unit force along j
    lap = fd_laplacian(u, X0)
    gp = fd_grad(p, X0)
    assert np.allclose(MU * lap, gp.T, atol=1e-5 * np.abs(gp).max())
    assert np.all(np.abs(np.einsum("ddj->j", fd_grad(u, X0))) < 1e-7)


def test_stokes_double_layer_velocity_pressure_pair():
    """(u, p) = (321, 322) for a source normal n solve mu lap u = grad p."""
    n = np.array([0.3, -0.5, 0.81])
    n = n / np.linalg.norm(n)
    u = lambda x: O.kernel_matrix(321, [MU], np.zeros((1, 3)), n[None, :], np.atleast_2d(x))
    p = lambda x: O.kernel_matrix(322, [MU], np.zeros((1, 3)), n[None, :], np.atleast_2d(x))[0]
    lap = fd_laplacian(u, X0, h=2e-3)
    gp = fd_grad(p, X0, h=1e-4)
    assert np.allclose(MU * lap, gp.T, atol=2e-4 * np.abs(gp).max())


def test_navier_double_layer_solves_lame_and_jump():
    """521 solves the Lame equation in the target and its integral over the unit
    sphere with outward normals is a constant multiple of the identity inside
    (magnitude 1, the double-layer jump) and zero outside."""
    lam = 2 * MU * NU / (1 - 2 * NU)
    n = np.array([-0.4, 0.1, 0.9])
    n = n / np.linalg.norm(n)
    u = lambda x: K(521, x, n)
    lap = fd_laplacian(u, X0, h=2e-3)
    div = lambda x: np.einsum("ddj->j", fd_grad(u, x, h=1e-4))
    gd = fd_grad(div, X0, h=1e-3)
    res = MU * lap + (lam + MU) * gd
    assert np.abs(res).max() < 5e-4 * np.abs(MU * lap).max()
    p, w = sphere_quadrature(20000)
    tot_in = None
    for x, inside in [(np.array([0.1, -0.2, 0.15]), True), (np.array([1.6, 0.3, -0.4]), False)]:
        m = O.kernel_matrix(521, [MU, NU], p, p, x[None, :])
        tot = sum(m[:, 3 * j:3 * j + 3] * w[j] for j in range(len(w)))
        if inside:
            c = tot[0, 0]
            assert abs(abs(c) - 1.0) < 2e-3 and np.allclose(tot, c * np.eye(3), atol=2e-3)
        else:
            assert np.allclose(tot, 0.0, atol=2e-3)
