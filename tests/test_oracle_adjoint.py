"""Pins the FP64 discrete adjoint of the oracle (oracle/adjoint_numpy.py) against
central finite differences of the oracle's own forward pieces, in FP64, on a small
strongly nonlinear case (Kerr + Raman, photo-ionisation loss, plasma absorption
and defocusing all active, rho up to ~1e18 cm^-3), and on an STE case (self-trapped
excitons on: the coupled (n_e, n_s) recurrence with STE re-ionisation, trapping and
STE avalanche, and the STE index term; rho and rho_s ~2e19 cm^-3). The rescaling
n_e + n_s <= n_a, which a physical field seldom reaches, is pinned on a synthetic
saturating input of the recurrence itself (three quarters of its samples rescaled).

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


STE = dict(enable_ste=True, tau_r=60e-15, tau_ste=150e-15)


def _case(nz=4, energy_uJ=40.0, dz=10e-6, **over):
    kw = dict(MATERIAL, nz=nz, Nt=64, N=17, R_factor=4.0, begin=0.0, end=nz * dz,
              t_window=400e-15, energy_uJ=energy_uJ, w0=30e-6, save_stride=1000,
              rho_t_stride=0, write_output=False, verbose=False)
    kw.update(over)
    cfg = O.Config(**kw)
    it = O.Integrator(cfg, envelope=O.envelope_lens_noise(noise=0.05, seed=7))
    return A.Adjoint(it)


@pytest.fixture(scope="module")
def adj():
    return _case()


@pytest.fixture(scope="module")
def adj_ste():
    return _case(energy_uJ=20.0, **STE)


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
    rho, rho_s = adj.plasma(u)
    assert rho.max() > 1e16 and not rho_s.any()
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
    rho, _ = adj.plasma(u)
    v = adj.lin.half_linear(u, adj.dz)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(v)
    w = _rand_c(u.shape, rng)
    Jd = _fd(lambda x: adj.alpha_half(x, rho), v, d, 1e-6)
    gv, grho = adj.alpha_half_adj(v, rho, w)
    assert abs(_ip(w, Jd) - _ip(gv, d)) <= 1e-6 * abs(_ip(w, Jd))
    dr = rng.standard_normal(rho.shape) * (rho + 1e10)
    Jr = _fd(lambda r: adj.alpha_half(v, r), rho, dr, 1e-6)
    assert abs(_ip(w, Jr) - float(np.sum(grho * dr))) <= 1e-6 * abs(_ip(w, Jr))


def _check_rhs_adjoint(adj, seed):
    rng = np.random.default_rng(seed)
    u = adj.it.u
    rho, rho_s = adj.plasma(u)
    y = adj.alpha_half(adj.lin.half_linear(u, adj.dz), rho)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(y)
    w = _rand_c(u.shape, rng)
    f = lambda x: adj.nl.split(x, rho, rho_s)[0]  # noqa: E731
    Jd = _fd(f, y, d, 1e-6)
    gy, grho, grho_s = adj.rhs_adj(y, rho, rho_s, w)
    lhs = _ip(w, Jd)
    assert abs(lhs - _ip(gy, d)) <= 1e-6 * abs(lhs)
    dr = rng.standard_normal(rho.shape) * (rho + 1e10)
    Jr = _fd(lambda r: adj.nl.split(y, r, rho_s)[0], rho, dr, 1e-6)
    assert abs(_ip(w, Jr) - float(np.sum(grho * dr))) <= 1e-6 * abs(_ip(w, Jr))
    return rng, y, rho, rho_s, w, grho_s


def test_rhs_adjoint(adj):
    _, _, _, _, _, grho_s = _check_rhs_adjoint(adj, 3)
    assert not grho_s.any()


def test_rhs_adjoint_each_term_matters(adj):
    """Dropping any one term of rhs_adj breaks the identity: the test above has teeth."""
    rng = np.random.default_rng(4)
    u = adj.it.u
    rho, rho_s = adj.plasma(u)
    y = adj.alpha_half(adj.lin.half_linear(u, adj.dz), rho)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(y)
    w = _rand_c(u.shape, rng)
    Jd = _fd(lambda x: adj.nl.split(x, rho, rho_s)[0], y, d, 1e-6)
    lhs = _ip(w, Jd)
    nl = adj.nl
    for attr in ("en_kerr_raman", "en_pi_loss", "en_plasma_absorb"):
        old = getattr(nl, attr)
        setattr(nl, attr, False)
        try:
            gy, _, _ = adj.rhs_adj(y, rho, rho_s, w)
        finally:
            setattr(nl, attr, old)
        assert abs(lhs - _ip(gy, d)) > 1e-4 * abs(lhs), attr


def test_plasma_adjoint(adj):
    rng = np.random.default_rng(5)
    u = adj.it.u
    rho, rho_s = adj.plasma(u)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(u)
    gr = rng.standard_normal(rho.shape)
    Jd = _fd(lambda x: adj.plasma(x)[0], u, d, 1e-6)
    lhs = float(np.sum(gr * Jd))
    rhs = _ip(adj.plasma_adj(u, rho, rho_s, gr, np.zeros_like(gr)), d)
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
    assert adj.plasma(adj.run(A.lg_field(theta, g, cfg, P), 2))[0].max() > 1e14
    assert abs(L - loss(theta)) <= 1e-12 * L
    grad = A.lg_field_grad(theta, g, cfg, P, g0)
    fd = np.array([(loss(theta + 1e-5 * e) - loss(theta - 1e-5 * e)) / 2e-5 for e in np.eye(2 * P)])
    err = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
    assert err <= 1e-5, (grad, fd)


# ---------------------------------------------------------------------- STE on
def test_ste_case_is_active(adj_ste):
    u = adj_ste.it.u
    nl = adj_ste.nl
    rho, rho_s = adj_ste.plasma(u)
    assert rho.max() > 1e19 and rho_s.max() > 1e19
    assert nl.ste_pref > 0 and nl.avalanche_s > 0 and nl.inv_tau_ste > 0
    assert nl.ste_pref * rho_s.max() * adj_ste.dz > 0.02            # STE index phase / step


def test_rhs_adjoint_ste(adj_ste):
    rng, y, rho, rho_s, w, grho_s = _check_rhs_adjoint(adj_ste, 13)
    nl = adj_ste.nl
    dr = rng.standard_normal(rho.shape) * (rho_s + 1e10)
    Jr = _fd(lambda r: nl.split(y, rho, r)[0], rho_s, dr, 1e-6)
    lhs = _ip(w, Jr)
    assert abs(lhs - float(np.sum(grho_s * dr))) <= 1e-6 * abs(lhs)
    # the STE index term carries weight: dropping it from the adjoint breaks the identity
    d = _rand_c(y.shape, rng, 1e-2) * np.abs(y)
    Jd = _fd(lambda x: nl.split(x, rho, rho_s)[0], y, d, 1e-6)
    old = nl.ste_pref
    nl.ste_pref = 0.0
    try:
        gy, _, _ = adj_ste.rhs_adj(y, rho, rho_s, w)
    finally:
        nl.ste_pref = old
    assert abs(_ip(w, Jd) - _ip(gy, d)) > 1e-4 * abs(_ip(w, Jd))


def test_plasma_adjoint_ste(adj_ste):
    """Both densities' gradients through the coupled recurrence and both LUTs."""
    rng = np.random.default_rng(15)
    u = adj_ste.it.u
    rho, rho_s = adj_ste.plasma(u)
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(u)
    ge = rng.standard_normal(rho.shape)
    gs = rng.standard_normal(rho.shape)
    Je = _fd(lambda x: adj_ste.plasma(x)[0], u, d, 1e-6)
    Js = _fd(lambda x: adj_ste.plasma(x)[1], u, d, 1e-6)
    for a, b in ((ge, 0 * gs), (0 * ge, gs), (ge, gs)):
        lhs = float(np.sum(a * Je) + np.sum(b * Js))
        rhs = _ip(adj_ste.plasma_adj(u, rho, rho_s, a, b), d)
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_recurrence_adjoint_rescaled_branch(adj_ste):
    """Synthetic saturating rates (n_a 1e17 cm^-3): most steps go through the
    n_e + n_s <= n_a rescaling; the adjoint of the whole recurrence with respect to
    I, W and Ws matches finite differences, and taking the rescaling as the identity
    fails the same check."""
    adj = _case(energy_uJ=20.0, rho_max=1e17, **STE)
    nl = adj.nl
    rng = np.random.default_rng(0)
    shp = (6, 64)
    I = 1e13 * rng.uniform(0.2, 1.0, shp)
    W = 1e30 * rng.uniform(0.2, 1.0, shp)
    Ws = 1e30 * rng.uniform(0.2, 1.0, shp)

    def fwd(I_, W_, Ws_):
        ne, ns = np.zeros(shp), np.zeros(shp)
        O.rate_recurrence_rows(I_, W_, Ws_, ne, ns, adj.dt, nl.avalanche, nl.avalanche_s,
                               nl.rho_max, nl.inv_taur, nl.inv_tau_ste, 1)
        return ne, ns
    ne, ns = fwd(I, W, Ws)
    assert np.mean(ne + ns >= nl.rho_max * (1 - 1e-12)) > 0.5
    dI, dW, dWs = (rng.standard_normal(shp) * x for x in (I, W, Ws))
    ge, gs = rng.standard_normal(shp), rng.standard_normal(shp)
    h = 1e-6
    p = fwd(I + h * dI, W + h * dW, Ws + h * dWs)
    m = fwd(I - h * dI, W - h * dW, Ws - h * dWs)
    lhs = float(np.sum(ge * (p[0] - m[0]) + gs * (p[1] - m[1])) / (2 * h))
    for rj, bound in ((True, 1e-6), (False, None)):
        gI, gW, gWs = adj.recurrence_adj(I, W, Ws, ne, ns, ge, gs, rescale_jacobian=rj)
        err = abs(lhs - float(np.sum(gI * dI + gW * dW + gWs * dWs))) / abs(lhs)
        if bound:
            assert err <= bound
        else:
            assert err > 1e-2


def test_step_adjoint_ste(adj_ste):
    rng = np.random.default_rng(16)
    u = adj_ste.it.u
    d = _rand_c(u.shape, rng, 1e-2) * np.abs(u)
    w = _rand_c(u.shape, rng) * adj_ste.mask_r
    Jd = _fd(adj_ste.step, u, d, 1e-6)
    lhs = _ip(w, Jd)
    assert abs(lhs - _ip(adj_ste.step_adj(u, w), d)) <= 1e-5 * abs(lhs)


def test_lg_gradient_ste_matches_finite_differences():
    adj = _case(nz=3, energy_uJ=20.0, dz=3e-6, **STE)
    g, cfg = adj.g, adj.cfg
    P = 4
    rng = np.random.default_rng(18)
    theta = np.concatenate([[1.0], 0.1 * rng.standard_normal(P - 1), 0.3 * rng.standard_normal(P)])
    hidden = theta + np.concatenate([0.05 * rng.standard_normal(P), 0.2 * rng.standard_normal(P)])
    F_t = A.fluence(adj.run(A.lg_field(hidden, g, cfg, P), 3), g)

    def loss(th):
        return A.fluence_loss(adj.run(A.lg_field(th, g, cfg, P), 3), g, F_t)[0]
    L, g0 = adj.gradient(A.lg_field(theta, g, cfg, P), 3, F_t)
    assert np.isfinite(L) and L > 1e-6
    assert adj.plasma(adj.run(A.lg_field(theta, g, cfg, P), 2))[1].max() > 1e18
    grad = A.lg_field_grad(theta, g, cfg, P, g0)
    fd = np.array([(loss(theta + 1e-5 * e) - loss(theta - 1e-5 * e)) / 2e-5 for e in np.eye(2 * P)])
    err = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
    assert err <= 1e-5, (grad, fd)
