"""0-D re-integration of the carrier densities over saved intensity histories, on the GPU.

Drop-in for ``Params`` and ``integre_0d`` of ``D3/sim/figures_article.py:55-108, 161-194``
(D3 = ``.github/deliverable (3)``). The reference integrates one history I(t) per call
in a Python loop, and ``fig13_electron_density_vs_z`` (``:266-303``) calls it once per
saved z plane and model variant. Here every history of a batch is one row of a single
kernel launch (``fp_rate0d_batch``): a thread per row runs the same exponential-step
recursion in FP64, with the Keldysh rates evaluated on the device from the PCHIP
coefficients (log10 I knots) that ``Params`` builds on the host exactly as the reference
does.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
from scipy.constants import c, epsilon_0, m_e
from scipy.interpolate import PchipInterpolator

from . import _lib
from ._lib import call, ptr, stream_ptr
from .keldysh import KeldyshSiO2, n_sellmeier

Q_E = 1.602176634e-19
HBAR = 1.054571817e-34


@dataclass
class Params:
    """Material / probe parameters of the post-processing model (figures_article.py:55-108).

    Same fields, defaults and derived quantities as the reference. The Keldysh rates
    of the valence band (Ui) and of the STE level (Us) are tabulated on 700 intensities
    log-spaced over [1, 1e17] W/cm^2 and interpolated with PCHIP in log10 I.
    """
    wavelength: float = 1030e-9
    Ui_eV: float = 9.0
    Us_eV: float = 6.0
    meff_rel: float = 0.64
    tau_c: float = 1.7e-15
    tau_r: float = 330e-15
    rho_max: float = 2.1e22
    n2: float = 3.54e-20
    lambda_probe: float = 490e-9
    E_tr_eV: float = 4.2
    probe_fwhm_fs: float = 26.3

    def __post_init__(self):
        self.n0 = n_sellmeier(self.wavelength)
        self.omega0 = 2 * np.pi * c / self.wavelength
        self.komega = 2 * np.pi / self.wavelength * self.n0
        self.Ui, self.Us = self.Ui_eV * Q_E, self.Us_eV * Q_E
        self.meff = self.meff_rel * m_e
        # inverse-Bremsstrahlung cross-section [cm^2] (Drude mass m_e), as build_grids
        self.sigmaomega = ((self.komega * Q_E**2 * self.tau_c) /
                           (self.n0**2 * m_e * epsilon_0 * self.omega0 *
                            (1.0 + (self.omega0 * self.tau_c)**2))) * 1e4
        self.beta_g = self.sigmaomega / self.Ui
        self.beta_s = self.sigmaomega / self.Us
        self.omega_p = 2 * np.pi * c / self.lambda_probe
        self.n0_probe = n_sellmeier(self.lambda_probe)
        self.nc_probe = (epsilon_0 * m_e * self.omega_p**2 / Q_E**2) * 1e-6
        w_tr = self.E_tr_eV * Q_E / HBAR
        self.f_ste = self.omega_p**2 / (w_tr**2 - self.omega_p**2)
        Igrid = np.logspace(0.0, 17.0, 700)
        Wg = np.maximum(np.nan_to_num(
            KeldyshSiO2(self.wavelength, self.Ui_eV, self.meff, self.n0).rate(Igrid)), 0.0)
        Ws = np.maximum(np.nan_to_num(
            KeldyshSiO2(self.wavelength, self.Us_eV, self.meff, self.n0).rate(Igrid)), 0.0)
        self._lI = np.log10(Igrid)
        self._Wg = PchipInterpolator(self._lI, Wg)
        self._Ws = PchipInterpolator(self._lI, Ws)
        self._Imin, self._Imax = Igrid[0], Igrid[-1]
        self._dev = {}

    def W_vb(self, I_Wcm2):
        """Valence-band Keldysh rate (host NumPy, as the reference)."""
        return np.maximum(self._Wg(np.log10(np.clip(I_Wcm2, self._Imin, self._Imax))), 0.0)

    def W_ste(self, I_Wcm2):
        """STE-level Keldysh rate (host NumPy, as the reference)."""
        return np.maximum(self._Ws(np.log10(np.clip(I_Wcm2, self._Imin, self._Imax))), 0.0)

    def device_tables(self, device):
        """(knots, coef_g, coef_s) float64 on `device`: the PCHIP breakpoints in log10 I
        and the 4 x (n-1) power-basis coefficients (scipy PPoly order), uploaded once."""
        key = str(device)
        if key not in self._dev:
            def up(a):
                return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
            self._dev[key] = (up(self._Wg.x), up(self._Wg.c), up(self._Ws.c))
        return self._dev[key]


def integre_0d_batch(t_fs, I_rows, prm: Params, avalanche=True, piegeage=True, ste=True,
                     device=None):
    """Re-integrate d rho_e/dt, d rho_s/dt over every row of `I_rows` (n_rows x Nt, W/cm^2)
    on the common time axis `t_fs` (fs). Returns (NE, NS), n_rows x Nt float64 torch
    tensors on the device [cm^-3] (NS is zero when ``ste`` is False).

    Model variants as the reference's flags: ``avalanche`` (sigma/Ui I rho depletion
    term), ``piegeage`` (trapping, -rho/tau_r and the STE feed), ``ste`` (the second
    population and its re-ionisation)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.type != "cuda":
        raise ValueError("integre_0d_batch runs on the HIP device")
    t = torch.as_tensor(np.asarray(t_fs, dtype=np.float64) * 1e-15, device=dev)
    I = torch.as_tensor(I_rows, dtype=torch.float64, device=dev)
    if I.dim() == 1:
        I = I[None, :]
    I = I.contiguous()
    n_rows, Nt = I.shape
    if t.numel() != Nt:
        raise ValueError(f"t_fs has {t.numel()} samples, the histories {Nt}")
    knots, cg, cs = prm.device_tables(dev)
    ne = torch.empty((n_rows, Nt), dtype=torch.float64, device=dev)
    ns = torch.empty_like(ne)
    p = _lib.Rate0DParams(n_rows, Nt, float(prm.rho_max), 1.0 / prm.tau_r, float(prm.beta_g),
                          float(prm.beta_s), int(bool(avalanche)), int(bool(piegeage)),
                          int(bool(ste)), int(knots.numel()), float(prm._Imin), float(prm._Imax))
    with torch.cuda.device(dev):
        call("fp_rate0d_batch", ptr(t), ptr(I), ptr(ne), ptr(ns), p, ptr(knots), ptr(cg),
             ptr(cs), stream_ptr())
    return ne, ns


def integre_0d(t_fs, I_Wcm2, prm: Params, avalanche=True, piegeage=True, ste=True):
    """Same signature and result as the reference (one history, NumPy in and out)."""
    ne, ns = integre_0d_batch(t_fs, np.asarray(I_Wcm2, dtype=np.float64)[None, :], prm,
                              avalanche, piegeage, ste)
    return ne[0].cpu().numpy(), ns[0].cpu().numpy()
