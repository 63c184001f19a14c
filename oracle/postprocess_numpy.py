"""NumPy FP64 restatement of the two secondary models next to the solver.

TEST INFRASTRUCTURE ONLY -- the oracle. Imported by ``tests/``; never by the
product package. ``D3`` below is ``.github/deliverable (3)``.

* ``Params0D`` / ``integre_0d_rows``: the 0-D carrier re-integration of
  ``D3/sim/figures_article.py:55-108, 161-194`` (Params, _phi1, _step, integre_0d),
  vectorised over a batch of intensity histories (the reference loops over one history
  per call in pure Python; the per-row arithmetic below is the same, element by element).
* ``shock_rhs`` / ``propagate_shock_rows``: the 1-D SPM + shock model of
  ``D3/sim/self_steepening.py:60-105``, vectorised over a batch of pulses.

Pins (tests/test_oracle_postprocess.py), chosen so that a dropped term, a wrong sign
or a wrong factor fails one of them:

* 0-D: (a) constant intensity, no depletion (rho_max huge), trapping only: the
  exponential step is exact for constant S, L, so n_e(t) = W tau (1 - e^{-t/tau}) to
  rounding; (b) multiphoton only with depletion: the recursion n_{k+1} = n_k (1 - W dt/n_a)
  + W dt has the closed form n_a (1 - (1 - W dt/n_a)^k); (c) avalanche with constant I,
  no depletion, no trapping: n_e = W/(beta I) (e^{beta I t} - 1); (d) the STE branch
  (no closed form): input/output vectors of the reference's own integre_0d
  (tests/golden/rate0d_ste.npz, written by tests/golden/make_rate0d_golden.py) for the
  three trapping / avalanche variants, which reach the n_e + n_s = n_a rescaling;
  (e) n_e, n_s >= 0 and n_e + n_s <= n_a on random histories.
* 1-D: (a) SPM only: E(z) = E0 exp(i K0 n2 |E0|^2 z) (RK4 error below 1e-10 at the
  step sizes used); (b) energy sum |E|^2 conserved by the shock term (the continuous
  model conserves it exactly; RK4 to 1e-9); (c) intensity follows the Burgers
  characteristics: max|dI/dt| reaches 3x its input value at (2/3) z_shock of the closed
  form; (d) the shock term moves the intensity peak to later times (sign).
"""
from __future__ import annotations

import numpy as np
from scipy.constants import c, epsilon_0, m_e
from scipy.interpolate import PchipInterpolator

from .reference_numpy import Keldysh, n_sellmeier

Q_E = 1.602176634e-19
HBAR = 1.054571817e-34


class Params0D:
    """figures_article.py:55-108 (the fields the 0-D model reads)."""

    def __init__(self, wavelength=1030e-9, Ui_eV=9.0, Us_eV=6.0, meff_rel=0.64, tau_c=1.7e-15,
                 tau_r=330e-15, rho_max=2.1e22):
        self.wavelength, self.Ui_eV, self.Us_eV = wavelength, Ui_eV, Us_eV
        self.meff_rel, self.tau_c, self.tau_r, self.rho_max = meff_rel, tau_c, tau_r, rho_max
        self.n0 = n_sellmeier(wavelength)                                  # :85
        self.omega0 = 2 * np.pi * c / wavelength
        self.komega = 2 * np.pi / wavelength * self.n0
        self.Ui, self.Us = Ui_eV * Q_E, Us_eV * Q_E
        self.meff = meff_rel * m_e
        self.sigmaomega = ((self.komega * Q_E**2 * tau_c) /                 # :92-94
                           (self.n0**2 * m_e * epsilon_0 * self.omega0 *
                            (1.0 + (self.omega0 * tau_c)**2))) * 1e4
        self.beta_g = self.sigmaomega / self.Ui                             # :95-96
        self.beta_s = self.sigmaomega / self.Us
        Igrid = np.logspace(0.0, 17.0, 700)                                 # :102-108
        Wg = np.maximum(np.nan_to_num(Keldysh(wavelength, Ui_eV, self.meff, self.n0).rate(Igrid)), 0.0)
        Ws = np.maximum(np.nan_to_num(Keldysh(wavelength, Us_eV, self.meff, self.n0).rate(Igrid)), 0.0)
        self.lI = np.log10(Igrid)
        self.Wg_spline = PchipInterpolator(self.lI, Wg)
        self.Ws_spline = PchipInterpolator(self.lI, Ws)
        self.Imin, self.Imax = Igrid[0], Igrid[-1]

    def W_vb(self, I):                                                      # :110-111
        return np.maximum(self.Wg_spline(np.log10(np.clip(I, self.Imin, self.Imax))), 0.0)

    def W_ste(self, I):                                                     # :113-114
        return np.maximum(self.Ws_spline(np.log10(np.clip(I, self.Imin, self.Imax))), 0.0)


def _phi1(x):
    """figures_article.py:161-162, elementwise."""
    x = np.asarray(x, dtype=np.float64)
    big = np.abs(x) > 1e-6
    safe = np.where(big, x, 1.0)
    return np.where(big, np.expm1(safe) / safe, 1.0 + x * (0.5 + x * (1.0 / 6.0 + x / 24.0)))


def _step(x, S, L, dt):
    """figures_article.py:164-166, elementwise: exact exponential step, 0 if not finite or < 0."""
    with np.errstate(over="ignore", invalid="ignore"):
        xn = np.exp(L * dt) * x + S * dt * _phi1(L * dt)
    return np.where(np.isfinite(xn) & (xn >= 0.0), xn, 0.0)


def integre_0d_rows(t_fs, I_rows, prm, avalanche=True, piegeage=True, ste=True, W=None, Ws=None):
    """figures_article.py:168-194 for every row of I_rows (n_rows x Nt) at once.

    W, Ws (optional, same shape as I_rows) replace the Keldysh rates; by default they are
    prm.W_vb(I), prm.W_ste(I) as in the reference."""
    t = np.asarray(t_fs, float) * 1e-15
    I = np.clip(np.atleast_2d(np.asarray(I_rows, float)), 0.0, None)
    W = prm.W_vb(I) if W is None else np.atleast_2d(W)
    Ws = (prm.W_ste(I) if ste else np.zeros_like(I)) if Ws is None else np.atleast_2d(Ws)
    na, inv_tr = prm.rho_max, 1.0 / prm.tau_r
    n_rows, Nt = I.shape
    ne = np.zeros(n_rows)
    ns = np.zeros(n_rows)
    NE, NS = np.zeros_like(I), np.zeros_like(I)
    for k in range(Nt - 1):
        dt = t[k + 1] - t[k]
        Iavg = 0.5 * (I[:, k] + I[:, k + 1])
        Wavg = 0.5 * (W[:, k] + W[:, k + 1])
        Wsav = 0.5 * (Ws[:, k] + Ws[:, k + 1])
        depl = np.minimum(np.maximum(1.0 - ne / na, 0.0), 1.0)
        Se = Wavg * depl + ((Wsav + prm.beta_s * Iavg * ne) * (ns / na) if ste else 0.0)
        Le = (prm.beta_g * Iavg * depl if avalanche else 0.0) - (inv_tr if piegeage else 0.0)
        Ss = inv_tr * ne if (ste and piegeage) else np.zeros(n_rows)
        Ls = -(Wsav + prm.beta_s * Iavg * ne) / na if ste else np.zeros(n_rows)
        ne_n = _step(ne, Se, Le, dt)
        ns_n = _step(ns, Ss, Ls, dt) if ste else np.zeros(n_rows)
        tot = ne_n + ns_n
        over = tot > na
        f = np.where(over, na / np.where(over, tot, 1.0), 1.0)
        ne_n = np.where(over, ne_n * f, ne_n)
        ns_n = np.where(over, ns_n * f, ns_n)
        ne, ns = ne_n, ns_n
        NE[:, k + 1], NS[:, k + 1] = ne, ns
    return NE, NS


# --------------------------------------------------------------------------
# 1-D SPM + shock (D3/sim/self_steepening.py:60-105)
# --------------------------------------------------------------------------
def shock_rhs(E, w, K0, n2, shock):
    """self_steepening.py:65-77 for rows E (n_rows x N); w (n_rows x N or N), shock
    (bool per row)."""
    I = np.abs(E) ** 2
    IE = I * E
    out = 1j * K0 * n2 * IE
    d = np.fft.ifft(1j * w * np.fft.fft(IE, axis=-1), axis=-1)             # :60-62
    return np.where(np.asarray(shock)[:, None], out - (n2 / c) * d, out)


def propagate_shock_rows(t_rows, E0_rows, z_max, nz, lam0=800e-9, n2=3.54e-20, shock=True,
                         nsave=None):
    """self_steepening.py:80-105 for a batch of pulses (one time grid and z_max per row).
    Returns (z n_rows x n_saves, E n_rows x n_saves x N)."""
    E = np.atleast_2d(np.asarray(E0_rows, dtype=np.complex128)).copy()
    n_rows, N = E.shape
    t_rows = np.broadcast_to(np.atleast_2d(np.asarray(t_rows, float)), (n_rows, N))
    dt = t_rows[:, 1] - t_rows[:, 0]
    w = np.stack([2 * np.pi * np.fft.fftfreq(N, d) for d in dt])
    K0 = 2 * np.pi / lam0
    dz = (np.broadcast_to(np.asarray(z_max, float), (n_rows,)) / nz)[:, None]
    sk = np.broadcast_to(np.asarray(shock, bool), (n_rows,))
    nsave = nsave or min(nz, 240)
    stride = max(1, nz // nsave)
    zs, Es = [np.zeros(n_rows)], [E.copy()]
    for i in range(nz):
        k1 = shock_rhs(E, w, K0, n2, sk)
        k2 = shock_rhs(E + 0.5 * dz * k1, w, K0, n2, sk)
        k3 = shock_rhs(E + 0.5 * dz * k2, w, K0, n2, sk)
        k4 = shock_rhs(E + dz * k3, w, K0, n2, sk)
        E = E + (dz / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)
        if (i + 1) % stride == 0:
            zs.append((i + 1) * dz[:, 0])
            Es.append(E.copy())
    return np.stack(zs, axis=1), np.stack(Es, axis=1)


def shock_distance_theory(t_p_s, I_p, n2=3.54e-20):
    """self_steepening.py:108-116."""
    return 1.0 / (3.0 * np.sqrt(2.0) * np.exp(-0.5)) * c * t_p_s / (n2 * I_p)


def max_slope_vs_z(Es, dt):
    """self_steepening.py:263-266."""
    return np.max(np.abs(np.gradient(np.abs(Es) ** 2, dt, axis=-1)), axis=-1)
