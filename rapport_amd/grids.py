"""Grids and operator tables of the propagator, built once on the host in FP64.

Drop-in for ``build_grids`` / ``envelope_gaussian_focused`` / ``ENVELOPES`` of
``D3/sim/grids.py:29-185`` (D3 = ``.github/deliverable (3)``): the returned dict has
the reference's keys with NumPy (host) arrays. The per-step kernels only need
1-D vectors along t, the Hankel matrix and a few tables; ``device_tables`` in
``operators.py`` uploads those. Envelopes are evaluated on the host in FP64 and
uploaded once.

What is computed (reference line in brackets):
* time grid ``tlist = linspace(-tmax, tmax, Nt, endpoint=False)``, ``tmax = 5 tp``
  or ``t_window/2`` [:32-36];
* quasi-discrete Hankel transform ``Y[i, j] = (2/j_N) J0(j_i j_j/j_N) / J1(j_j)^2``
  on the first N-1 Bessel zeros, ``r_m = j_m R / j_N``, ``rho_m = j_m / R`` [:38-44];
* Sellmeier ``delta_k(w) = k(w) - k0 - k1 Omega`` with w clipped to the Sellmeier band
  and k1 by central difference dw = w0/1000 [:46-60];
* spectral band mask, self-steepening ``T``, space-time focusing ``U``, ``1/U`` and the
  floored ``mask/max(U, 0.1)`` [:62-111];
* Raman response ``R_f = fft(ifftshift(R(t))) dt`` [:113-116];
* Drude cross-section, avalanche / recombination / STE coefficients, ``invE2``,
  STE index prefactor and the radial absorber ``exp(-(r/0.9R)^20)`` [:121-138];
* Keldysh rates on 17 decades x 900 knots, their PCHIP in linear I (field loss) and
  the 4096-point log10 LUTs (carrier kernel) [:140-161].
"""
from __future__ import annotations

import numpy as np
from scipy.constants import c, epsilon_0, m_e
from scipy.constants import elementary_charge as q_e
from scipy.interpolate import PchipInterpolator
from scipy.special import j0, j1, jn_zeros

from .keldysh import KeldyshSiO2, sellmeier_n2m1

HBAR = 1.054571817e-34
NLUT = 4096


def intensity_knots():
    """I = 10^k (1 + 0.01 j), k = 0..16, j = 0..899 (W/cm^2), increasing."""
    return np.ravel(10 ** (np.arange(17) + 0.0)[:, None] * (0.01 * np.arange(900) + 1)[None, :])


def hankel_matrix(N):
    """(Y, zeros): Y of order N-1 with column J1^2 normalisation (not symmetric)."""
    jz = jn_zeros(0, N)
    jm = jz[:N - 1]
    jN = jz[N - 1]
    Y = (2.0 / jN) * j0(np.outer(jm, jm) / jN) / (j1(jm) ** 2)[None, :]
    return Y, jz


def _k_of(w):
    lam_um = 2 * np.pi * c / w * 1e6
    return (w / c) * np.sqrt(1.0 + sellmeier_n2m1(lam_um))


def build_grids(cfg) -> dict:
    komega, omega0, Nt, N = cfg.komega, cfg.omega0, cfg.Nt, cfg.N
    tp = cfg.delta_t / np.sqrt(2 * np.log(2))
    tmax = 5 * tp if getattr(cfg, "t_window", None) is None else 0.5 * cfg.t_window
    dt = float(2 * tmax / Nt)
    tlist = np.linspace(-tmax, tmax, Nt, endpoint=False)
    ff = np.fft.fftfreq(Nt, d=dt)
    Om = 2 * np.pi * ff

    Y, jz = hankel_matrix(N)
    R = cfg.R_factor * cfg.w0
    jN = jz[N - 1]
    rlist = jz[:N - 1] * R / jN
    rholist = jz[:N - 1] / R

    w_lo, w_hi = 2 * np.pi * c / 5.0e-6, 2 * np.pi * c / 0.18e-6
    w_safe = np.clip(omega0 + Om, w_lo, w_hi)
    dw = omega0 / 1000.0
    k1 = (_k_of(omega0 + dw) - _k_of(omega0 - dw)) / (2 * dw)
    delta_k = _k_of(w_safe) - float(komega) - k1 * Om

    nu = (cfg.frequency + ff) / cfg.frequency
    if cfg.enable_spectral_filter:
        nu_lo, nu_hi, edge = (c / 5.0e-6) / cfg.frequency, (c / 0.18e-6) / cfg.frequency, 0.05
        spec_mask = 0.25 * (1.0 + np.tanh((nu - nu_lo) / edge)) * (1.0 + np.tanh((nu_hi - nu) / edge))
    else:
        spec_mask = np.ones(Nt)
    T_op = (1.0 + ff / cfg.frequency) * spec_mask if cfg.enable_self_steepening else np.ones(Nt)
    U_op = 1.0 + k1 * Om / komega if cfg.enable_space_time_focusing else np.ones(Nt)
    inv_U_op = 1.0 / U_op
    inv_U_nl = spec_mask / np.maximum(U_op, 0.1) if cfg.enable_space_time_focusing else np.ones(Nt)

    R_t = np.zeros(Nt)
    pos = tlist > 0
    R_t[pos] = cfg.Omega_r2 * cfg.tau_s * np.exp(-tlist[pos] / cfg.tau_d) * np.sin(tlist[pos] / cfg.tau_s)
    R_f = np.fft.fft(np.fft.ifftshift(R_t)) * dt

    tt, rr = np.meshgrid(tlist, rlist)
    _, rhorho = np.meshgrid(tlist, rholist)

    sigma = ((float(komega) * q_e ** 2 * cfg.tau_c)
             / (cfg.n0 ** 2 * cfg.meff_drude * epsilon_0 * omega0 * (1.0 + (omega0 * cfg.tau_c) ** 2))) * 1e4
    aval = sigma / cfg.Ui if cfg.enable_avalanche else 0.0
    aval_s = sigma / cfg.Us if (cfg.enable_ste and cfg.enable_avalanche) else 0.0
    inv_taur = 1.0 / cfg.tau_r if cfg.enable_recombination else 0.0
    inv_tau_ste = 1.0 / cfg.tau_ste if (cfg.tau_ste and cfg.enable_ste) else 0.0
    invE2 = 0.5 * cfg.n0 * c * epsilon_0 * 1e-4

    rho_c = (epsilon_0 * m_e * omega0 ** 2 / q_e ** 2) * 1e-6
    w_tr = cfg.E_tr_eV * q_e / HBAR
    f_ste = omega0 ** 2 / (w_tr ** 2 - omega0 ** 2)
    ste_pref = ((omega0 / c) * f_ste / (2.0 * cfg.n0 * rho_c)
                if (cfg.enable_ste and cfg.enable_ste_index) else 0.0)
    mask_r_1d = np.exp(-(rlist / (0.9 * R)) ** 20)
    mask_r = np.repeat(mask_r_1d[:, None], Nt, axis=1)

    I_k = intensity_knots()
    kopt = dict(monotone=cfg.keldysh_monotone, beta_den=cfg.keldysh_beta_den)
    W_g = np.maximum(np.nan_to_num(
        KeldyshSiO2(cfg.wavelength, cfg.Ui_eV, cfg.meff, cfg.n0, **kopt).rate(I_k)), 0.0)
    W_s = np.maximum(np.nan_to_num(
        KeldyshSiO2(cfg.wavelength, cfg.Us_eV, cfg.meff, cfg.n0, **kopt).rate(I_k)), 0.0)
    sp_g = PchipInterpolator(I_k, W_g)
    sp_s = PchipInterpolator(I_k, W_s)
    lmin, lmax = float(np.log10(I_k[0])), float(np.log10(I_k[-1]))
    logI = np.linspace(lmin, lmax, NLUT)
    kel = dict(
        f_spline=sp_g,
        W_LUT=np.nan_to_num(np.abs(sp_g(10.0 ** logI)), nan=0.0, posinf=0.0, neginf=0.0),
        W_LUT_s=np.nan_to_num(np.abs(sp_s(10.0 ** logI)), nan=0.0, posinf=0.0, neginf=0.0),
        NLUT=NLUT, logImin=lmin, inv_dlog=float((NLUT - 1) / (lmax - lmin)),
        Imin=float(I_k[0]), Imax=float(I_k[-1]),
        knots=I_k,
    )
    return dict(
        komega=float(komega), tmax=tmax, dt=dt, tlist=tlist, ff=ff,
        Y=Y, R=R, j1last=jN, rlist=rlist, rholist=rholist,
        delta_k=delta_k, k1=k1, T_op=T_op, inv_U_op=inv_U_op, inv_U_nl=inv_U_nl, R_f=R_f,
        spec_mask=spec_mask, rr=rr, tt=tt, rhorho=rhorho,
        sigmaomega=sigma, avalanche_coef=aval, avalanche_coef_s=aval_s,
        inv_taur_eff=inv_taur, inv_tau_ste=inv_tau_ste,
        ste_pref=ste_pref, f_ste=f_ste, invE2=invE2, mask_r=mask_r, mask_r_1d=mask_r_1d,
        b=cfg.b, E0=cfg.E0, keldysh=kel,
    )


def envelope_gaussian_focused(rr, tt, cfg, g):
    """Gaussian beam that focuses at z = 0, started at z = cfg.begin
    (q = 1 + 2 i begin / b), times exp(-t^2/tp^2) [grids.py:180-183]."""
    tp = cfg.delta_t / np.sqrt(2 * np.log(2))
    q = 1 + 1j * 2 * cfg.begin / g["b"]
    return (g["E0"] / q) * np.exp(-rr ** 2 / cfg.w0 ** 2 / q) * np.exp(-tt ** 2 / tp ** 2)


def seeded_noise(shape, seed):
    """xi = (x + i y)/sqrt(2) with x, y ~ N(0, 1) from numpy default_rng(seed),
    drawn as one float64 block of shape (Nr, Nt, 2)."""
    xy = np.random.default_rng(seed).standard_normal(tuple(shape) + (2,))
    return (xy[..., 0] + 1j * xy[..., 1]) / np.sqrt(2.0)


def envelope_lens_noise(lens_f=None, noise=0.0, seed=0):
    """Envelope factory: gaussian_focused x thin lens exp(-i k0 r^2 / (2 f)) (converging
    for f > 0, same sign convention as the focused Gaussian) x (1 + noise * xi)."""
    def envelope(rr, tt, cfg, g):
        u = envelope_gaussian_focused(rr, tt, cfg, g)
        if lens_f:
            u = u * np.exp(-1j * g["komega"] * rr ** 2 / (2.0 * lens_f))
        if noise:
            u = u * (1.0 + noise * seeded_noise(u.shape, seed))
        return u
    return envelope


ENVELOPES = {"gaussian_focused": envelope_gaussian_focused}
