"""Pins of the oracle for the two secondary models (0-D re-integration, 1-D SPM + shock)
and host-side checks of their product wrappers (no GPU)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import postprocess_numpy as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------------- 0-D model
@pytest.fixture(scope="module")
def prm():
    return O.Params0D()


def _const(n, v):
    return np.full((1, n), float(v))


def test_0d_trapping_exact_exponential(prm):
    """Constant W, no depletion, trapping only: n(t) = W tau (1 - e^{-t/tau}) exactly."""
    prm.rho_max, keep = 1e300, prm.rho_max
    try:
        t_fs = np.linspace(0.0, 2000.0, 1601)
        W0 = 3.7e30
        ne, ns = O.integre_0d_rows(t_fs, _const(t_fs.size, 1e12), prm, avalanche=False,
                                   piegeage=True, ste=False, W=_const(t_fs.size, W0))
        tau = prm.tau_r
        exact = W0 * tau * (-np.expm1(-t_fs * 1e-15 / tau))
        np.testing.assert_allclose(ne[0], exact, rtol=1e-12, atol=0)
        assert not ns.any()
    finally:
        prm.rho_max = keep


def test_0d_depletion_closed_form(prm):
    """Multiphoton only: n_{k+1} = n_k + W dt (1 - n_k/n_a) -> n_a (1 - (1 - W dt/n_a)^k)."""
    t_fs = np.arange(0.0, 1000.0, 0.5)
    W0 = 2.0e37                                        # W dt / n_a ~ 4.8e-4 per step
    ne, _ = O.integre_0d_rows(t_fs, _const(t_fs.size, 1e13), prm, avalanche=False,
                              piegeage=False, ste=False, W=_const(t_fs.size, W0))
    a = W0 * 0.5e-15 / prm.rho_max
    k = np.arange(t_fs.size)
    exact = prm.rho_max * -np.expm1(k * np.log1p(-a))
    np.testing.assert_allclose(ne[0], exact, rtol=1e-10, atol=1e-6 * prm.rho_max * 1e-10)
    assert ne[0, -1] > 0.5 * prm.rho_max               # depletion really acts


def test_0d_avalanche_exact_exponential(prm):
    """Constant I and W, no depletion, no trapping: n = W/(beta I) (e^{beta I t} - 1)."""
    prm.rho_max, keep = 1e300, prm.rho_max
    try:
        t_fs = np.linspace(0.0, 300.0, 1201)
        I0, W0 = 4e13, 1.0e28
        ne, _ = O.integre_0d_rows(t_fs, _const(t_fs.size, I0), prm, avalanche=True,
                                  piegeage=False, ste=False, W=_const(t_fs.size, W0))
        g = prm.beta_g * I0
        exact = W0 / g * np.expm1(g * t_fs * 1e-15)
        assert g * t_fs[-1] * 1e-15 > 1.0                # the avalanche term matters
        np.testing.assert_allclose(ne[0], exact, rtol=1e-11, atol=0)
    finally:
        prm.rho_max = keep


def test_0d_ste_branch_matches_reference_vectors(prm):
    g = np.load(os.path.join(HERE, "golden", "rate0d_ste.npz"))
    for name, (av, pg) in {"full": (True, True), "noav": (False, True),
                           "notrap": (True, False)}.items():
        ne, ns = O.integre_0d_rows(g["t_fs"], g["I"], prm, av, pg, True)
        np.testing.assert_allclose(ne, g["ne_" + name], rtol=1e-13, atol=0)
        np.testing.assert_allclose(ns, g["ns_" + name], rtol=1e-13, atol=0)
    # the vectors exercise the n_e + n_s <= n_a rescaling and a live STE population
    assert np.max(g["ne_full"] + g["ns_full"]) >= prm.rho_max * (1 - 1e-12)
    assert np.max(g["ns_full"]) > 1e20


def test_0d_invariants_random(prm):
    rng = np.random.default_rng(1)
    t_fs = np.sort(rng.uniform(-200, 200, 500))        # non-uniform grid
    I = 10 ** rng.uniform(10, 14.8, (20, 500))
    for av in (False, True):
        for pg in (False, True):
            for ste in (False, True):
                ne, ns = O.integre_0d_rows(t_fs, I, prm, av, pg, ste)
                assert (ne >= 0).all() and (ns >= 0).all()
                assert (ne + ns <= prm.rho_max * (1 + 1e-15)).all()
                if not ste:
                    assert not ns.any()


def test_product_params_tables_equal_oracle(prm):
    """The product's Params builds the same coefficients the kernel reads."""
    from rapport_amd.reintegrate import Params
    p = Params()
    for a in ("n0", "sigmaomega", "beta_g", "beta_s"):
        assert getattr(p, a) == getattr(prm, a), a
    np.testing.assert_array_equal(p._Wg.x, prm.Wg_spline.x)
    np.testing.assert_array_equal(p._Wg.c, prm.Wg_spline.c)
    np.testing.assert_array_equal(p._Ws.c, prm.Ws_spline.c)
    I = 10 ** np.linspace(-1, 18, 333)
    np.testing.assert_array_equal(p.W_vb(I), prm.W_vb(I))
    np.testing.assert_array_equal(p.W_ste(I), prm.W_ste(I))


# ------------------------------------------------------------------------- 1-D model
def _pulse(nt, tp_fs, Ip_Wcm2=5e13):
    t_s = np.linspace(-8 * tp_fs, 8 * tp_fs, nt) * 1e-15
    E0 = np.sqrt(Ip_Wcm2 * 1e4) * np.exp(-(t_s / (tp_fs * 1e-15)) ** 2 / 2.0)
    return t_s, E0


def test_shock_spm_only_closed_form():
    t_s, E0 = _pulse(256, 30.0)
    lam0, n2 = 800e-9, 3.54e-20
    z_max, nz = 1e-4, 3000                             # RK4 phase error ~ 1e-9
    z, E = O.propagate_shock_rows(t_s, E0, z_max, nz, lam0, n2, shock=False, nsave=10)
    K0 = 2 * np.pi / lam0
    exact = E0[None, :] * np.exp(1j * K0 * n2 * np.abs(E0[None, :]) ** 2 * z[0][:, None])
    assert K0 * n2 * np.max(np.abs(E0)) ** 2 * z_max > 10.0     # many radians of SPM
    np.testing.assert_allclose(E[0], exact, rtol=0, atol=1e-8 * np.max(np.abs(E0)))


def test_shock_energy_conserved_and_peak_moves_later():
    t_s, E0 = _pulse(1024, 30.0)
    n2, Ip = 3.54e-20, 5e13 * 1e4
    zs = O.shock_distance_theory(30e-15, Ip, n2)
    z, E = O.propagate_shock_rows(t_s, E0, 0.5 * zs, 1600, n2=n2, shock=True, nsave=4)
    en = np.sum(np.abs(E[0]) ** 2, axis=-1)
    np.testing.assert_allclose(en, en[0], rtol=1e-10)     # RK4 drift ~ dz^4: 2e-11 here
    I = np.abs(E[0]) ** 2
    assert t_s[np.argmax(I[-1])] > t_s[np.argmax(I[0])] + 2 * (t_s[1] - t_s[0])


def test_shock_burgers_distance():
    """max|dI/dt| reaches 3x its input value at (1 - 1/3) z_shock (Burgers law)."""
    tp = 30.0
    t_s, E0 = _pulse(1024, tp)
    n2, Ip = 3.54e-20, 5e13 * 1e4
    zs = O.shock_distance_theory(tp * 1e-15, Ip, n2)
    z, E = O.propagate_shock_rows(t_s, E0, 0.9 * zs, 1200, n2=n2, shock=True, nsave=400)
    dt = t_s[1] - t_s[0]
    slope = O.max_slope_vs_z(E[0], dt)
    idx = int(np.argmax(slope >= 3 * slope[0]))
    zc = z[0, idx - 1] + (z[0, idx] - z[0, idx - 1]) * (3 * slope[0] - slope[idx - 1]) / (
        slope[idx] - slope[idx - 1])
    assert abs(zc / zs - 2.0 / 3.0) < 0.02


def test_shock_save_plan_matches_reference_rule():
    from rapport_amd.shock import save_plan
    for nz, nsave in [(4000, 400), (4000, None), (7, None), (100, 30), (5, 10)]:
        ns = nsave or min(nz, 240)
        stride = max(1, nz // ns)
        count = 1 + sum(1 for i in range(nz) if (i + 1) % stride == 0)
        assert save_plan(nz, nsave) == (stride, count)


def test_entry_points_reject_bad_arguments():
    from rapport_amd import _lib
    p = _lib.Rate0DParams(4, 0, 2.1e22, 1.0, 1.0, 1.0, 1, 1, 0, 700, 1.0, 1e17)
    dummy = ctypes.c_void_p(0x1000)
    with pytest.raises(RuntimeError, match="Nt >= 1"):
        _lib.call("fp_rate0d_batch", dummy, dummy, dummy, dummy, ctypes.byref(p), dummy, dummy,
                  None, None)
    p.Nt, p.ste = 10, 1
    with pytest.raises(RuntimeError, match="STE needs coef_s"):
        _lib.call("fp_rate0d_batch", dummy, dummy, dummy, dummy, ctypes.byref(p), dummy, dummy,
                  None, None)
    with pytest.raises(RuntimeError, match="N must be"):
        _lib.call("fp_shock_propagate", dummy, dummy, dummy, dummy, 1, 3000, 10, 1, dummy, None)
    with pytest.raises(RuntimeError, match="needs work"):
        _lib.call("fp_shock_propagate", dummy, dummy, None, dummy, 1, 8192, 10, 1, dummy, None)
    with pytest.raises(RuntimeError, match="stride"):
        _lib.call("fp_shock_propagate", dummy, dummy, dummy, dummy, 1, 1024, 10, 0, dummy, None)
