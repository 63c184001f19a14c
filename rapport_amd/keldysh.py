"""Sellmeier index and Keldysh photo-ionisation rate of fused silica (host, FP64).

Drop-in for ``D3/sim/keldysh.py`` (D3 = ``.github/deliverable (3)``): ``n_sellmeier``
(:25-29), ``KeldyshSiO2`` (:35-148), ``validate`` (:154-246) and the two
asymptotic limits (:266-305). These build the rate tables once per run on the
host; the per-step evaluation happens on the GPU from the uploaded tables.

Keldysh rate (Couairon et al., PRB 71, 125435 (2005), Eqs. (7)-(8)), in cm^-3 s^-1:

    W = (2 w/9 pi) (w m / (hbar sqrt(G)))^{3/2} Q(g, x) exp(-a <x+1>)
    g = w sqrt(m U)/(e E),  G = g^2/(1+g^2),  X = 1/(1+g^2)
    a = pi (K(G) - E(G)) / E(X)
    x = (2/pi) (U/hbar w) E(X)/sqrt(G),  nu = <x+1> - x
    Q = sqrt(pi / (2 K(X))) sum_n exp(-n a) Dawson(pi sqrt((n + 2 nu)/(beta_den K(X) E(X))))

with beta_den = 4 (the paper's convention, the reference default).
"""
from __future__ import annotations

import numpy as np
from scipy.constants import c, epsilon_0
from scipy.constants import elementary_charge as q_e
from scipy.interpolate import PchipInterpolator
from scipy.special import dawsn, ellipe, ellipk

HBAR = 1.054571817e-34
M_E = 9.1093837015e-31

# Malitson (1965) three-term Sellmeier coefficients for fused silica; wavelengths in um.
SELLMEIER_B = np.array([0.6961663, 0.4079426, 0.8974794])
SELLMEIER_L2 = np.array([0.0684043, 0.1162414, 9.896161]) ** 2


def sellmeier_n2m1(lam_um):
    """n^2 - 1 at wavelength(s) lam_um (micrometres)."""
    l2 = np.asarray(lam_um, dtype=np.float64) ** 2
    return sum(B * l2 / (l2 - L2) for B, L2 in zip(SELLMEIER_B, SELLMEIER_L2))


def n_sellmeier(lam_m):
    return float(np.sqrt(1.0 + sellmeier_n2m1(lam_m * 1e6)))


class KeldyshSiO2:
    """Keldysh photo-ionisation rate W(I) [cm^-3 s^-1] for intensity I [W/cm^2]."""

    def __init__(self, wavelength, Ui_eV, meff, n0, N_sum=60, monotone=False, beta_den=4.0,
                 lut_logI=(0.0, 17.0), lut_points=6000):
        self.U = Ui_eV * q_e
        self.meff = meff
        self.n0 = n0
        self.omega = 2 * np.pi * c / wavelength
        self.hbar = HBAR
        self.N_sum = N_sum
        self.monotone = monotone
        self.beta_den = float(beta_den)
        self._lut_logI, self._lut_points = lut_logI, lut_points
        self._mono = None

    def gamma(self, I_Wcm2):
        E = np.sqrt(2.0 * np.asarray(I_Wcm2, dtype=np.float64) * 1e4 / (self.n0 * c * epsilon_0))
        return np.maximum(self.omega * np.sqrt(self.meff * self.U) / (q_e * np.maximum(E, 1e-300)),
                          1e-12)

    def rate_raw(self, intensity_Wcm2):
        g = self.gamma(intensity_Wcm2)
        g2 = g * g
        G, X = g2 / (1.0 + g2), 1.0 / (1.0 + g2)
        sG = np.sqrt(G)
        KG, EG, KX, EX = ellipk(G), ellipe(G), ellipk(X), ellipe(X)
        x = (2.0 * self.U * EX) / (np.pi * sG * HBAR * self.omega)
        n_photon = np.floor(x + 1.0)
        nu = n_photon - x
        a = np.pi * (KG - EG) / EX
        n = np.arange(self.N_sum).reshape((-1,) + (1,) * np.ndim(x))
        arg = np.pi * np.sqrt(np.maximum(n + 2.0 * nu, 0.0) / (self.beta_den * KX * EX))
        Q = np.sqrt(np.pi / (2.0 * KX)) * np.sum(np.exp(-n * a) * dawsn(arg), axis=0)
        pref = (2.0 * self.omega / (9.0 * np.pi)) * (self.omega * self.meff / (HBAR * sG)) ** 1.5
        W = pref * Q * np.exp(-a * n_photon) * 1e-6
        return np.nan_to_num(W, nan=0.0, posinf=0.0, neginf=0.0)

    def _monotone_spline(self):
        if self._mono is None:
            lI = np.linspace(self._lut_logI[0], self._lut_logI[1], self._lut_points)
            W = np.maximum.accumulate(np.maximum(self.rate_raw(10.0 ** lI), 0.0))
            self._mono = (lI, PchipInterpolator(lI, W))
        return self._mono

    def rate(self, intensity_Wcm2, monotone=None):
        mono = self.monotone if monotone is None else monotone
        if not mono:
            return self.rate_raw(intensity_Wcm2)
        lI, sp = self._monotone_spline()
        I = np.asarray(intensity_Wcm2, dtype=np.float64)
        lq = np.log10(np.clip(I, 10.0 ** lI[0], 10.0 ** lI[-1]))
        return np.maximum(np.nan_to_num(sp(lq), nan=0.0, posinf=0.0, neginf=0.0), 0.0)


def validate(wavelength=800e-9, Ui_eV=9.0, meff_rel=0.64, sigma6=9.6e-70, rho_at=2.1e22,
             verbose=True):
    """Checks the rate against Couairon 2005: multiphoton normalisation and order,
    gamma = 1 at 3.5e13 W/cm^2, and the beta -> 1, exp(-a) -> e^2/(16 g^2)
    asymptotes that fix the summation convention. Raises AssertionError."""
    n0 = n_sellmeier(wavelength)
    kel = KeldyshSiO2(wavelength, Ui_eV, meff_rel * M_E, n0)
    out = {}
    gb = 100.0
    G, X = gb * gb / (1 + gb * gb), 1.0 / (1 + gb * gb)
    beta_inf = np.pi ** 2 / (4.0 * ellipk(X) * ellipe(X))
    a_inf = np.pi * (ellipk(G) - ellipe(G)) / ellipe(X)
    out["beta_mpi_limit"] = float(beta_inf)
    out["exp_alpha_ratio"] = float(np.exp(-a_inf) / (np.e ** 2 / (16 * gb * gb)))
    I_lo = np.array([5e11, 1e12, 2e12])
    ratio = kel.rate(I_lo) / (sigma6 * I_lo ** 6 * rho_at)
    order = np.diff(np.log(kel.rate(I_lo))) / np.diff(np.log(I_lo))
    out["mpi_ratio"], out["order_low_I"] = ratio, order
    I = np.logspace(11.5, 14.2, 600)
    slope = np.diff(np.log(np.maximum(kel.rate(I), 1e-300))) / np.diff(np.log(I))
    out["n_decreasing"] = int((slope < -1e-9).sum())
    out["slope_min"] = float(slope.min())
    sr = np.diff(np.log(np.maximum(kel.rate_raw(I), 1e-300))) / np.diff(np.log(I))
    out["n_decreasing_raw"] = int((sr < 0).sum())
    out["slope_min_raw"] = float(sr.min())
    out["gamma_at_3.5e13"] = float(kel.gamma(3.5e13))
    out["W_3.5e13_beta4"] = float(kel.rate(np.array([3.5e13]))[0])
    kel2 = KeldyshSiO2(wavelength, Ui_eV, meff_rel * M_E, n0, beta_den=2.0)
    out["W_3.5e13_beta2"] = float(kel2.rate(np.array([3.5e13]))[0])
    out["beta_den"] = kel.beta_den
    if verbose:
        print(f"beta(gamma=100) {beta_inf:.5f}  exp(-a)/(e^2/16g^2) {out['exp_alpha_ratio']:.5f}")
        print(f"W/W_MPI {ratio[0]:.3f}/{ratio[1]:.3f}/{ratio[2]:.3f}  order {order.mean():.2f}  "
              f"gamma(3.5e13) {out['gamma_at_3.5e13']:.3f}")
    assert 0.8 < ratio[1] < 1.25, f"rate at 1e12 W/cm^2 is {ratio[1]:.3f} x sigma6 I^6 rho_at"
    assert 5.5 < order.mean() < 6.5, f"low-intensity log-log slope {order.mean():.2f}, expected ~6"
    assert 0.95 < out["gamma_at_3.5e13"] < 1.05, \
        f"Keldysh parameter at 3.5e13 W/cm^2 is {out['gamma_at_3.5e13']:.3f}, expected ~1"
    assert abs(beta_inf - 1.0) < 1e-3, f"Dawson-argument factor at gamma = 100 is {beta_inf:.5f}"
    assert abs(out["exp_alpha_ratio"] - 1.0) < 1e-3, \
        f"exp(-alpha) / (e^2 / 16 gamma^2) at gamma = 100 is {out['exp_alpha_ratio']:.5f}"
    return out


def keldysh_multiphoton(I_Wcm2, wavelength=800e-9, Ui_eV=9.0, meff=None, n0=None):
    """gamma >> 1 limit: (2w/9pi)(w m/hbar)^{3/2} Dawson(sqrt(2 nu)) e^{2<x+1>(1-1/4g^2)} (16 g^2)^{-<x+1>}."""
    meff = 0.64 * M_E if meff is None else meff
    n0 = n_sellmeier(wavelength) if n0 is None else n0
    w, U = 2 * np.pi * c / wavelength, Ui_eV * q_e
    g = KeldyshSiO2(wavelength, Ui_eV, meff, n0).gamma(I_Wcm2)
    xt = (U / (HBAR * w)) * (1.0 + 1.0 / (4.0 * g ** 2))
    mx = np.floor(xt + 1.0)
    W = ((2.0 * w / (9.0 * np.pi)) * (w * meff / HBAR) ** 1.5 * dawsn(np.sqrt(2.0 * (mx - xt)))
         * np.exp(2.0 * mx * (1.0 - 1.0 / (4.0 * g ** 2))) * (1.0 / (16.0 * g ** 2)) ** mx)
    return np.nan_to_num(W, nan=0.0, posinf=0.0, neginf=0.0) * 1e-6


def keldysh_tunnel(I_Wcm2, wavelength=800e-9, Ui_eV=9.0, meff=None, n0=None):
    """gamma << 1 limit: (2/9pi^2)(mU/hbar^2)^{3/2}(U/hbar)(hbar w/(U g))^{5/2}
    exp[-(pi/2) sqrt(mU) U/(hbar e E)(1 - g^2/8)]."""
    meff = 0.64 * M_E if meff is None else meff
    n0 = n_sellmeier(wavelength) if n0 is None else n0
    w, U = 2 * np.pi * c / wavelength, Ui_eV * q_e
    E = np.sqrt(2.0 * np.asarray(I_Wcm2, float) * 1e4 / (n0 * c * epsilon_0))
    g = KeldyshSiO2(wavelength, Ui_eV, meff, n0).gamma(I_Wcm2)
    W = ((2.0 / (9.0 * np.pi ** 2)) * (meff * U / HBAR ** 2) ** 1.5 * (U / HBAR)
         * (HBAR * w / (U * g)) ** 2.5
         * np.exp(-(np.pi / 2.0) * np.sqrt(meff * U) * U / (HBAR * q_e * np.maximum(E, 1e-300))
                  * (1.0 - g ** 2 / 8.0)))
    return np.nan_to_num(W, nan=0.0, posinf=0.0, neginf=0.0) * 1e-6


if __name__ == "__main__":
    validate()
