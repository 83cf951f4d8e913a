"""Pins of the CPU oracle's pieces by mathematics that fixes them
independently of the reference's code: textbook definitions, closed forms,
invariants and exact special cases.  Each check is chosen so that a
plausible slip (wrong coefficient, sign, index or operand order) fails it.
"""
import ctypes

import numpy as np
import pytest

from hypar_amd import decks
from oracle.oracle import Oracle, lib

DP = ctypes.POINTER(ctypes.c_double)
GAMMA = 1.4


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(DP)


# ---------------------------------------------------------------- WENO5-JS
def _beta_from_definition(f):
    """Jiang-Shu smoothness indicators from their integral definition
    beta_k = sum_{l=1,2} int_{cell} h^(2l-1) (d^l p_k/dx^l)^2 dx, where p_k is
    the quadratic whose cell averages over stencil k equal the point values
    (finite-difference WENO treats f_j as cell averages of the primitive's
    derivative).  Evaluated with exact polynomial algebra on h = 1."""
    P = np.polynomial.polynomial
    betas = []
    # stencils relative to the target cell j (f index 2): k=1 {-2,-1,0}, k=2 {-1,0,1}, k=3 {0,1,2}
    for offs in ((-2, -1, 0), (-1, 0, 1), (0, 1, 2)):
        # cell averages of a + b x + c x^2 over [o-1/2, o+1/2] = a + b o + c (o^2 + 1/12)
        A = np.array([[1.0, o, o * o + 1.0 / 12.0] for o in offs])
        coef = np.linalg.solve(A, np.array([f[o + 2] for o in offs]))
        d1 = P.polyder(coef, 1)
        d2 = P.polyder(coef, 2)
        b = 0.0
        for d in (d1, d2):
            sq = P.polymul(d, d)
            integ = P.polyint(sq)
            b += P.polyval(0.5, integ) - P.polyval(-0.5, integ)
        betas.append(b)
    return np.array(betas)


def test_weno_weights_match_integral_definition():
    rng = np.random.default_rng(7)
    L = lib()
    eps = 1e-6
    c = np.array([0.1, 0.6, 0.3])
    for _ in range(200):
        f = rng.normal(size=5) * rng.choice([1e-3, 1e-1, 1.0, 10.0])
        out = (ctypes.c_double * 3)()
        L.orc_weno_js_weights(*[ctypes.c_double(v) for v in f], ctypes.c_double(eps), out)
        beta = _beta_from_definition(f)
        a = c / (beta + eps) ** 2
        ref = a / a.sum()
        assert np.allclose(np.array(out[:]), ref, rtol=1e-10, atol=1e-13), (f, out[:], ref)


def test_weno_constant_data_gives_optimal_weights():
    out = (ctypes.c_double * 3)()
    lib().orc_weno_js_weights(*[ctypes.c_double(2.5)] * 5, ctypes.c_double(1e-6), out)
    assert np.allclose(out[:], [0.1, 0.6, 0.3], rtol=0, atol=2e-16)


def test_weno_reconstruction_exact_for_quadratics():
    """Any convex weights reproduce the interface value of a quadratic
    primitive: f_j = cell averages of p, result = p(x_{j+1/2})."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b, cc = rng.normal(size=3)
        h = rng.uniform(0.05, 0.5)
        x = (np.arange(5) - 2) * h
        avg = a + b * x + cc * (x * x + h * h / 12.0)
        got = lib().orc_weno5_interp(_p(avg), 1e-6)
        xf = 0.5 * h
        assert abs(got - (a + b * xf + cc * xf * xf)) < 1e-13 * (1 + abs(a) + abs(b) + abs(cc))


def test_weno_fifth_order_on_smooth_data():
    """Error of the reconstruction of a smooth function away from critical
    points falls ~32x per halving of h (fifth order)."""
    errs = []
    for h in (0.2, 0.1, 0.05):
        x = (np.arange(5) - 2) * h
        # cell averages of g = sin(x+1) reconstruct g at x+h/2
        avg = (np.cos(x - h / 2 + 1) - np.cos(x + h / 2 + 1)) / h
        got = lib().orc_weno5_interp(_p(avg), 1e-6)
        errs.append(abs(got - np.sin(h / 2 + 1)))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > 4.0, (errs, rates)


# ------------------------------------------- mapped / Z / Yamaleev-Carpenter
WTYPES = {"js": 0, "mapped": 1, "borges": 2, "yc": 3, "optimal": 4}


def _weights(kind, f, eps=1e-6):
    out = (ctypes.c_double * 3)()
    lib().orc_weno_weights(WTYPES[kind], *[ctypes.c_double(v) for v in f], ctypes.c_double(eps), out)
    return np.array(out[:])


@pytest.mark.parametrize("kind", ["js", "mapped", "borges", "yc", "optimal"])
def test_weights_convex_and_optimal_on_constant_data(kind):
    c = np.array([0.1, 0.6, 0.3])
    assert np.allclose(_weights(kind, [2.5] * 5), c, rtol=0, atol=4e-16)
    rng = np.random.default_rng(11)
    for _ in range(100):
        w = _weights(kind, rng.normal(size=5) * rng.choice([1e-2, 1.0, 1e2]))
        assert abs(w.sum() - 1.0) < 1e-14 and (w >= 0).all()


def test_modified_weights_approach_optimal_faster_on_smooth_data():
    """On smooth data the Jiang-Shu weights deviate from the optimal ones by
    O(h^2); the mapping of Henrick et al. (g(c) = c, g'(c) = 0), the WENO-Z
    indicator tau = |b1-b3| = O(h^5) of Borges et al. and the squared
    4th difference of Yamaleev & Carpenter all cut that deviation by orders
    of magnitude."""
    c = np.array([0.1, 0.6, 0.3])
    h = 0.05
    x = (np.arange(5) - 2) * h
    avg = (np.cos(x - h / 2 + 1) - np.cos(x + h / 2 + 1)) / h
    dev = {k: np.abs(_weights(k, avg) - c).max() for k in WTYPES}
    assert dev["js"] > 1e-4
    for k in ("mapped", "borges", "yc"):
        assert dev[k] < 1e-2 * dev["js"], dev
    assert dev["optimal"] == 0.0  # no_limiting: the weights are c on any data


# ------------------------------------------------------------- eigen system
def _euler_flux(u, d):
    rho = u[0]
    v = u[1:4] / rho
    P = (u[4] - 0.5 * rho * v.dot(v)) * (GAMMA - 1)
    f = rho * v[d] * np.array([1.0, v[0], v[1], v[2], 0.0])
    f[1 + d] += P
    f[4] = (u[4] + P) * v[d]
    return f


def _state(rng):
    rho = rng.uniform(0.5, 2.0)
    v = rng.normal(size=3) * 0.5
    P = rng.uniform(0.3, 2.0)
    return np.array([rho, rho * v[0], rho * v[1], rho * v[2], P / (GAMMA - 1) + 0.5 * rho * v.dot(v)])


def _LR(u, d):
    L = np.zeros(25)
    R = np.zeros(25)
    lib().orc_left_eigenvectors(_p(u), L.ctypes.data_as(DP), GAMMA, d)
    lib().orc_right_eigenvectors(_p(u), R.ctypes.data_as(DP), GAMMA, d)
    return L.reshape(5, 5), R.reshape(5, 5)


def _eigvals(u, d):
    rho = u[0]
    v = u[1:4] / rho
    P = (u[4] - 0.5 * rho * v.dot(v)) * (GAMMA - 1)
    c = np.sqrt(GAMMA * P / rho)
    lam = np.full(5, v[d])
    lam[1 + d] = v[d] - c
    lam[4] = v[d] + c
    return lam


@pytest.mark.parametrize("d", [0, 1, 2])
def test_eigenvectors_diagonalise_flux_jacobian(d):
    rng = np.random.default_rng(11 + d)
    for _ in range(20):
        u = _state(rng)
        L, R = _LR(u, d)
        assert np.allclose(L @ R, np.eye(5), atol=1e-13)
        # Jacobian by central differences of the textbook Euler flux
        J = np.zeros((5, 5))
        for k in range(5):
            e = np.zeros(5)
            e[k] = 1e-6 * max(1.0, abs(u[k]))
            J[:, k] = (_euler_flux(u + e, d) - _euler_flux(u - e, d)) / (2 * e[k])
        A = R @ np.diag(_eigvals(u, d)) @ L
        assert np.allclose(A, J, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_roe_average_property(d):
    """A(u_roe) (uR - uL) = f(uR) - f(uL) for the ideal-gas Roe average."""
    rng = np.random.default_rng(5 + d)
    for _ in range(20):
        uL, uR = _state(rng), _state(rng)
        ua = np.zeros(5)
        lib().orc_roe_average(ua.ctypes.data_as(DP), _p(uL), _p(uR), GAMMA)
        L, R = _LR(ua, d)
        A = R @ np.diag(_eigvals(ua, d)) @ L
        lhs = A @ (uR - uL)
        rhs = _euler_flux(uR, d) - _euler_flux(uL, d)
        assert np.allclose(lhs, rhs, rtol=1e-11, atol=1e-12)


# ---------------------------------------------- first derivative (ref test)
def test_fourth_order_derivative_exact_on_cubic():
    """Mirrors tests/FirstDerivative/test_first_derivative.c:126-171 of the
    reference: all stencils, biased ones included, are exact on cubics."""
    n, g = 24, 3
    h = 0.1
    x = (np.arange(n + 2 * g) - g) * h
    f = 1.0 + 2 * x - 3 * x ** 2 + 0.5 * x ** 3
    Df = np.zeros_like(f)
    lib().orc_first_derivative4_1d(n, g, _p(f), Df.ctypes.data_as(DP))
    exact = (2 - 6 * x + 1.5 * x ** 2) * h
    assert np.max(np.abs(Df - exact)) < 1e-10


# ------------------------------------------------------ operator invariants
SCHEMES = [("components", "rusanov"), ("components", "roe"),
           ("characteristic", "rusanov"), ("characteristic", "roe")]


@pytest.mark.parametrize("interp,upwind", SCHEMES)
def test_free_stream_preservation(interp, upwind):
    dims = (8, 10, 12)
    o = Oracle(dims, interp, upwind, Re=100.0, Minf=0.5)
    u = np.zeros((dims[2], dims[1], dims[0], 5))
    u[...] = _state(np.random.default_rng(1))
    assert np.max(np.abs(o.rhs(u))) < 1e-13


@pytest.mark.parametrize("interp,upwind", SCHEMES)
def test_inviscid_conservation_and_translation(interp, upwind):
    w = decks.config0(12, 1)
    w.dims = (12, 10, 14)
    u = decks.initial_state(w)
    o = Oracle(w.dims, interp, upwind)
    h = o.hyperbolic(u)
    # telescoping flux differences: sum of hyp/dxinv vanishes per direction sum
    dx = [decks.dxinv_1d(decks.grid_1d(n))[3:-3] for n in w.dims]
    vol = 1.0 / (dx[0][None, None, :] * dx[1][None, :, None] * dx[2][:, None, None])
    tot = (h * vol[..., None]).reshape(-1, 5).sum(axis=0)
    scale = (np.abs(h) * vol[..., None]).reshape(-1, 5).sum(axis=0)
    assert np.all(np.abs(tot) <= 1e-12 * scale + 1e-14)
    # shifting the periodic field shifts the operator (up to ulp-level dxinv)
    us = np.roll(u, 5, axis=2)
    hs = o.hyperbolic(us)
    assert np.max(np.abs(hs - np.roll(h, 5, axis=2))) < 1e-11 * np.max(np.abs(h))
