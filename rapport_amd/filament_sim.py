"""Public façade: drop-in for ``D3/sim/filament_sim.py`` (D3 = ``.github/deliverable (3)``).

``run()`` builds a ``Config`` and returns ``Integrator(cfg, envelope).propagate()``
(filament_sim.py:59-82), with the same keyword defaults as the reference's
``run()`` (which differ from ``Config``'s on purpose). ``FIELD_TOGGLES`` lists the
six switchable channels of the field equation (filament_sim.py:87-94).
"""
from .config import Config, code_fingerprint
from .grids import ENVELOPES, build_grids, envelope_gaussian_focused, envelope_lens_noise
from .integrator import Integrator
from .keldysh import (SELLMEIER_B, SELLMEIER_L2, KeldyshSiO2, keldysh_multiphoton,
                      keldysh_tunnel, n_sellmeier)
from .operators import LinearOperator, NonlinearOperator


def run(*, Nz, Nt, Nr, wavelength, energy_uJ=15.0, peak_power_W=None,
        envelope="gaussian_focused", w0=10e-6, delta_t=263e-15, begin=-600e-6, end=1800e-6,
        n2=3.54e-20, Ui_eV=9.0, R_factor=100.0, save_stride=1, ckpt_every=500, verbose=True,
        out_dir=None, **material):
    """Propagate one pulse. ``material`` may hold any Config field (the six
    field toggles, the carrier toggles, and the additions t_window /
    write_output / precision). ``Nr`` is the Hankel order: the field has Nr-1
    radial rows."""
    cfg = Config(nz=Nz, Nt=Nt, N=Nr, wavelength=wavelength, energy_uJ=energy_uJ,
                 peak_power_W=peak_power_W, w0=w0, delta_t=delta_t, begin=begin, end=end,
                 n2=n2, Ui_eV=Ui_eV, R_factor=R_factor, save_stride=save_stride,
                 ckpt_every=ckpt_every, verbose=verbose, out_dir=out_dir, **material)
    return Integrator(cfg, envelope=envelope).propagate()


FIELD_TOGGLES = [
    "enable_kerr_instantaneous",
    "enable_kerr_raman",
    "enable_self_steepening",
    "enable_photoionization_loss",
    "enable_plasma_defocusing",
    "enable_plasma_absorption",
]

__all__ = [
    "run", "FIELD_TOGGLES", "Config", "code_fingerprint",
    "build_grids", "envelope_gaussian_focused", "envelope_lens_noise", "ENVELOPES",
    "LinearOperator", "NonlinearOperator", "Integrator",
    "n_sellmeier", "KeldyshSiO2", "keldysh_multiphoton", "keldysh_tunnel",
    "SELLMEIER_B", "SELLMEIER_L2",
]
