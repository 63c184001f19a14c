"""Pins the FP64 discrete adjoint of the oracle (oracle/adjoint_numpy.py) against
central finite differences of the oracle's own forward pieces, in FP64, on a small
strongly nonlinear case (Kerr + Raman, photo-ionisation loss, plasma absorption
and defocusing all active, rho up to ~1e18 cm^-3).

For a piece y = f(x) and random directions d (input) and w (output cotangent),
the adjoint must satisfy Re<w, J d> = Re<f^H(w), d>, with J d from
(f(x + h d) - f(x - h d)) / 2h. Tolerances are relative to |Re<w, J d>|; they sit
above the finite-difference error (h^2 terms and the kinks of the clamps / the
log-LUT) and far below the size of any dropped term or sign slip.
"""
import numpy as np
import pytest
import scipy.special as sp

from oracle import adjoint_numpy as A
from oracle import reference_numpy as O
from tests._cases import MATERIAL


def _case(nz=4, energy_uJ=40.0, dz=10e-6):
    kw = dict(MATERIAL, nz=nz, Nt=64, N=17, R_factor=4.0, begin=0.0, end=nz * dz,
              t_window=400e-15, energy_uJ=energy_uJ, w0=30e-6, save_stride=1000,
              rho_t_stride=0, write_output=False, verbose=False)
    cfg = O.Config(**kw)
    it = O.Integrator(cfg, envelope=O.envelope_lens_noise(noise=0.05, seed=7))
    return A.Adjoint(it)


@pytest.fixture(scope="module")
def adj():
    return _case()


def _rand_c(shape, rng, scale=1.0):
    return scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _ip(a, b):
    return float(np.real(np.sum(np.conj(a) * b)))


def test_laguerre_closed_forms():
    x = np.linspace(0, 12, 50)
    np.testing.assert_allclose(A.laguerre(0, x), 1.0)
    np.testing.assert_allclose(A.laguerre(1, x), 1 - x)
    np.testing.assert_allclose(A.laguerre(2, x), 1 - 2 * x + x ** 2 / 2)
    np.testing.assert_allclose(A.laguerre(3, x), 1 - 3 * x + 1.5 * x ** 2 - x ** 3 / 6, atol=1e-12)
    for p in range(16):
        np.testing.assert_allclose(A.laguerre(p, x), sp.eval_laguerre(p, x), rtol=1e-10, atol=1e-10)


def test_case_is_strongly_nonlinear(adj):
    u = adj.it.u
    rho = adj.plasma(u)
    assert rho.max() > 1e16
    a2 = np.abs(u) ** 2
    photo, pa, _ = adj.photo_derivs(a2, rho)
    assert photo.max() > 0 and np.abs(pa).max() > 0


def test_half_linear_adjoint_exact(adj):
    rng = np.random.default_rng(1)
    v = _rand_c(adj.it.u.shape, rng)
    w = _rand_c(adj.it.u.shape, rng)
    lhs = _ip(w, adj.lin.half_linear(v, adj.dz))
    rhs = _ip(adj.half_linear_adj(w), v)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def _fd(f, x, d, h):
    return (f(x + h * d) - f(x - h * d)) / (2 * h)


def test_alpha_half_adjoint(adj):
    rng = np.random.default_rng(2)
    u = adj.it.u
    rho = adj.plasma(u)
    v = adj.lin.half_linear(u, adj.dz)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(v)
    w = _rand_c(u.shape, rng)
    Jd = _fd(lambda x: adj.alpha_half(x, rho), v, d, 1e-6)
    gv, grho = adj.alpha_half_adj(v, rho, w)
    assert abs(_ip(w, Jd) - _ip(gv, d)) <= 1e-6 * abs(_ip(w, Jd))
    dr = rng.standard_normal(rho.shape) * (rho + 1e10)
    Jr = _fd(lambda r: adj.alpha_half(v, r), rho, dr, 1e-6)
    assert abs(_ip(w, Jr) - float(np.sum(grho * dr))) <= 1e-6 * abs(_ip(w, Jr))


def test_rhs_adjoint(adj):
    rng = np.random.default_rng(3)
    u = adj.it.u
    rho = adj.plasma(u)
    y = adj.alpha_half(adj.lin.half_linear(u, adj.dz), rho)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(y)
    w = _rand_c(u.shape, rng)
    f = lambda x: adj.nl.split(x, rho)[0]  # noqa: E731
    Jd = _fd(f, y, d, 1e-6)
    gy, grho = adj.rhs_adj(y, rho, w)
    lhs = _ip(w, Jd)
    assert abs(lhs - _ip(gy, d)) <= 1e-6 * abs(lhs)
    dr = rng.standard_normal(rho.shape) * (rho + 1e10)
    Jr = _fd(lambda r: adj.nl.split(y, r)[0], rho, dr, 1e-6)
    assert abs(_ip(w, Jr) - float(np.sum(grho * dr))) <= 1e-6 * abs(_ip(w, Jr))


def test_rhs_adjoint_each_term_matters(adj):
    """Dropping any one term of rhs_adj breaks the identity: the test above has teeth."""
    rng = np.random.default_rng(4)
    u = adj.it.u
    rho = adj.plasma(u)
    y = adj.alpha_half(adj.lin.half_linear(u, adj.dz), rho)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(y)
    w = _rand_c(u.shape, rng)
    Jd = _fd(lambda x: adj.nl.split(x, rho)[0], y, d, 1e-6)
    lhs = _ip(w, Jd)
    nl = adj.nl
    for attr in ("en_kerr_raman", "en_pi_loss", "en_plasma_absorb"):
        old = getattr(nl, attr)
        setattr(nl, attr, False)
        try:
            gy, _ = adj.rhs_adj(y, rho, w)
        finally:
            setattr(nl, attr, old)
        assert abs(lhs - _ip(gy, d)) > 1e-4 * abs(lhs), attr


def test_plasma_adjoint(adj):
    rng = np.random.default_rng(5)
    u = adj.it.u
    rho = adj.plasma(u)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(u)
    gr = rng.standard_normal(rho.shape)
    Jd = _fd(adj.plasma, u, d, 1e-6)
    lhs = float(np.sum(gr * Jd))
    rhs = _ip(adj.plasma_adj(u, rho, gr), d)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_step_adjoint(adj):
    rng = np.random.default_rng(6)
    u = adj.it.u
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(u)
    w = _rand_c(u.shape, rng) * adj.mask_r
    Jd = _fd(adj.step, u, d, 1e-6)
    lhs = _ip(w, Jd)
    assert abs(lhs - _ip(adj.step_adj(u, w), d)) <= 1e-5 * abs(lhs)


def test_lg_gradient_matches_finite_differences():
    adj = _case(nz=3, energy_uJ=15.0, dz=3e-6)      # < 1 rad of Kerr phase per step
    g, cfg = adj.g, adj.cfg
    P = 4
    rng = np.random.default_rng(8)
    theta = np.concatenate([[1.0], 0.1 * rng.standard_normal(P - 1), 0.3 * rng.standard_normal(P)])
    hidden = theta + np.concatenate([0.05 * rng.standard_normal(P), 0.2 * rng.standard_normal(P)])
    F_t = A.fluence(adj.run(A.lg_field(hidden, g, cfg, P), 3), g)

    def loss(th):
        return A.fluence_loss(adj.run(A.lg_field(th, g, cfg, P), 3), g, F_t)[0]
    L, g0 = adj.gradient(A.lg_field(theta, g, cfg, P), 3, F_t)
    assert np.isfinite(L) and L > 1e-6
    assert adj.plasma(adj.run(A.lg_field(theta, g, cfg, P), 2)).max() > 1e14
    assert abs(L - loss(theta)) <= 1e-12 * L
    grad = A.lg_field_grad(theta, g, cfg, P, g0)
    fd = np.array([(loss(theta + 1e-5 * e) - loss(theta - 1e-5 * e)) / 2e-5 for e in np.eye(2 * P)])
    err = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
    assert err <= 1e-5, (grad, fd)
