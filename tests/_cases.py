"""Shared workload definitions and seeded inputs for the parity tests.

C1: 128 x 512 (N=129), R = 1 mm, 800 fs window, 2 uJ / 85 fs / w0 = 100 um,
    1 % seeded noise, dz = 10 um (200 steps = nz 199, end 1.99 mm), STE off.
C2: 1024 x 4096 (N=1025), R = 1.5 mm, 1.6 ps window, 5 uJ, lens f = 5 cm, 1 % noise,
    dz = 5 um (4000 steps = nz 3999), STE off.
Material: fused silica at 800 nm, n2 = 3.2e-20, Ui = 9 eV, rho_max = 2.1e22 cm^-3,
tau_c = 1.27 fs, tau_r = 150 fs.
"""
import numpy as np

MATERIAL = dict(wavelength=800e-9, n2=3.2e-20, Ui_eV=9.0, rho_max=2.1e22, tau_c=1.27e-15,
                tau_r=150e-15, delta_t=85e-15, w0=100e-6, enable_ste=False)

C1 = dict(nz=199, Nt=512, N=129, R_factor=10.0, begin=0.0, end=1.99e-3, t_window=800e-15,
          energy_uJ=2.0, lens_f=None, noise=0.01, seed=1)
C2 = dict(nz=3999, Nt=4096, N=1025, R_factor=15.0, begin=0.0, end=3999 * 5e-6, t_window=1.6e-12,
          energy_uJ=5.0, lens_f=0.05, noise=0.01, seed=2)


def cfg_kwargs(case, **over):
    kw = dict(MATERIAL)
    kw.update({k: v for k, v in case.items() if k not in ("lens_f", "noise", "seed")})
    kw.update(save_stride=max(1, case["nz"] + 1), rho_t_stride=0, write_output=False,
              verbose=False)
    kw.update(over)
    return kw


def envelope_args(case):
    return dict(lens_f=case["lens_f"], noise=case["noise"], seed=case["seed"])


def pulse_field(Nr, Nt, dt, invE2, peaks, seed=0, width_fs=60.0):
    """Rows of Gaussian pulses with the given peak intensities (W/cm^2) and 3 % noise."""
    rng = np.random.default_rng(seed)
    t = (np.arange(Nt) - Nt / 2) * dt
    peaks = np.asarray(peaks, dtype=np.float64).reshape(-1, 1)
    I = peaks * np.exp(-2 * t[None, :] ** 2 / (width_fs * 1e-15) ** 2)
    I = I * (1 + 0.03 * rng.standard_normal((Nr, Nt)))
    ph = rng.uniform(0, 2 * np.pi, (Nr, Nt))
    return np.sqrt(np.abs(I) / invE2) * np.exp(1j * ph)


def rel_l2(a, b):
    """||a - b|| / ||b|| in float64 / complex128 (float32 squares of 1e19-scale densities
    overflow)."""
    a, b = np.asarray(a), np.asarray(b)
    a = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)
    b = b.astype(np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
