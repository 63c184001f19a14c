"""Solver configuration: every tunable of the propagator and the derived constants.

Drop-in for ``Config`` / ``code_fingerprint`` of the reference
(``D3/sim/config.py:31-255``; D3 = ``.github/deliverable (3)``): same field names,
defaults and derived attributes, so keyword arguments that a reference caller
passes to ``run()`` / ``Config`` mean the same thing here. Additional fields
(all defaulting to the reference behaviour):

* ``t_window``  full width of the time window in s (reference: fixed 10 tp);
* ``write_output``  write result.npz / params.json (reference: always);
* ``precision``  "fp64" (drop-in arithmetic) or "mixed" (complex64 field,
  split FP16 Hankel GEMM, FP64 plasma scan and propagator phase);
* ``keldysh_monotone``, ``keldysh_beta_den``  the ``monotone`` and ``beta_den`` arguments
  of the Keldysh rate behind the photo-ionisation tables (reference: False, 4; its
  KeldyshSiO2 takes them, D3/sim/keldysh.py:80-148, but build_grids does not pass them).
"""
from __future__ import annotations

import datetime
import hashlib
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np
from scipy.constants import c, epsilon_0, m_e
from scipy.constants import elementary_charge as q_e

from .keldysh import n_sellmeier

_HERE = Path(__file__).resolve().parent

# Every file whose change alters results: the host solver modules and the
# HIP sources. A cached result.npz stamped with another fingerprint is stale.
_FINGERPRINT_FILES = ("config.py", "keldysh.py", "grids.py", "operators.py", "integrator.py",
                      "filament_sim.py", "csrc/common.hpp", "csrc/fft.hpp", "csrc/gemm_f64.hpp",
                      "csrc/nl.hpp", "csrc/plasma.hpp", "csrc/rapport_amd.hip")


def code_fingerprint() -> str:
    """12-hex-digit hash of the solver sources (as config.py:31-45 does for its own)."""
    h = hashlib.sha256()
    for name in _FINGERPRINT_FILES:
        f = _HERE / name
        if f.exists():
            h.update(f.read_bytes())
    return h.hexdigest()[:12]


@dataclass
class Config:
    # numerical grid
    nz: int = 40000
    Nt: int = 4096
    N: int = 1025
    begin: float = -600e-6
    end: float = 7400e-6
    R_factor: float = 40.0
    save_stride: int = 1
    ckpt_every: int = 500
    verbose: bool = True
    out_dir: Optional[str] = None
    # laser
    wavelength: float = 1030e-9
    energy_uJ: float = 15.0
    peak_power_W: Optional[float] = None
    w0: float = 10e-6
    delta_t: float = 263e-15
    z_focus_air_um: Optional[float] = None
    # material (SiO2)
    n2: float = 2.4e-20
    Ui_eV: float = 9.0
    meff_rel: float = 0.64
    meff_drude_rel: float = 1.0
    tau_c: float = 1.7e-15
    tau_r: float = 330e-15
    f_R: float = 0.18
    tau_d: float = 32e-15
    tau_s: float = 12e-15
    rho_max: float = 2.1e22
    # self-trapped excitons
    Us_eV: float = 6.0
    enable_ste: bool = True
    tau_ste: Optional[float] = None
    # field-equation channels
    enable_kerr_instantaneous: bool = True
    enable_kerr_raman: bool = True
    enable_self_steepening: bool = True
    enable_photoionization_loss: bool = True
    enable_plasma_defocusing: bool = True
    enable_plasma_absorption: bool = True
    enable_space_time_focusing: bool = True
    enable_spectral_filter: bool = True
    enable_ste_index: bool = True
    E_tr_eV: float = 4.2
    # carrier-equation channels
    enable_avalanche: bool = True
    enable_recombination: bool = True
    # probe
    lambda_probe: float = 490e-9
    # recording
    rho_t_stride: int = 15
    rho_r_stride: int = 1
    absorb_time_bins_fs: Optional[tuple] = None
    rho_snapshot_t_fs: Optional[float] = None
    # adaptive step (reserved; the fixed-step integrator ignores them, as in the reference)
    adaptive: bool = False
    dphi_max: float = 0.05
    drho_rel_max: float = 0.10
    dz_min: float = 10e-9
    dz_max: float = 200e-9
    dz_init: Optional[float] = None
    # additions
    t_window: Optional[float] = None
    write_output: bool = True
    precision: str = "fp64"
    keldysh_monotone: bool = False    # rate tables from KeldyshSiO2.rate(monotone=True)
    keldysh_beta_den: float = 4.0     # KeldyshSiO2 beta_den (the reference's default 4)

    def __post_init__(self):
        lam = self.wavelength
        self.omega0 = 2 * np.pi * c / lam
        self.omega_probe = 2 * np.pi * c / self.lambda_probe
        self.n0_probe = n_sellmeier(self.lambda_probe)
        self.nc_probe = (epsilon_0 * m_e * self.omega_probe ** 2 / q_e ** 2) * 1e-6
        self.frequency = c / lam
        self.n0 = n_sellmeier(lam)
        if self.z_focus_air_um is not None:
            # paraxial refraction at a flat face: focus distance x n0, waist unchanged
            self.begin = -self.n0 * self.z_focus_air_um * 1e-6
        self.Ui = self.Ui_eV * q_e
        self.Us = self.Us_eV * q_e
        self.meff = self.meff_rel * m_e
        self.meff_drude = self.meff_drude_rel * m_e
        self.chi3 = 4 / 3 * epsilon_0 * self.n0 ** 2 * c * self.n2
        self.Omega_r2 = 1.0 / self.tau_s ** 2 + 1.0 / self.tau_d ** 2
        self.komega = 2 * np.pi / lam * self.n0
        self.dz = (self.end - self.begin) / self.nz
        self.b = self.komega * self.w0 ** 2
        if self.dz_init is None:
            self.dz_init = min(self.dz, self.dz_max)
        # on-axis peak field from the pulse energy (Gaussian in r and t):
        # P = E / (tp sqrt(pi/2)), tp = FWHM / sqrt(2 ln 2); I0 = 2 P / (pi w0^2)
        if self.peak_power_W is not None:
            I0 = 2.0 * self.peak_power_W / (np.pi * self.w0 ** 2)
        else:
            tp = self.delta_t / np.sqrt(2 * np.log(2))
            I0 = 2.0 * (self.energy_uJ * 1e-6 / (tp * np.sqrt(np.pi / 2))) / (np.pi * self.w0 ** 2)
        self.I0_Wcm2 = I0 * 1e-4
        self.E0 = float(np.sqrt(2.0 * self.I0_Wcm2 * 1e4 / (self.n0 * c * epsilon_0)))
        if self.out_dir is None:
            self.out_dir = (f"sio2_{datetime.datetime.now():%Y%m%d_%H%M%S}"
                            f"_E{self.energy_uJ * 1e3:.0f}nJ_{self.delta_t * 1e15:.0f}fs")
        if self.precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {self.precision!r}")
        pts = (self.w0 * 1e6) / (self.R_factor * self.w0 * 1e6 / (self.N - 1))
        if pts < 3:
            print(f"[WARN] RADIAL UNDER-SAMPLING: {pts:.1f} pts/w0. Raise N.")
        elif pts < 6:
            print(f"[WARN] Radial sampling marginal: {pts:.1f} pts/w0.")
