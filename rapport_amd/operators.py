"""Split-step operators on the GPU: drop-in for ``LinearOperator`` /
``NonlinearOperator`` of ``D3/sim/operators.py:22-183`` (D3 = ``.github/deliverable (3)``).

Arrays are torch tensors on the current HIP device (complex128 fields,
float64 densities) instead of cupy arrays; the arithmetic runs in the
hand-written gfx950 kernels of ``librapport_amd.so`` through its C ABI.

* ``LinearOperator.half_linear(u, dz)``: FFT_t -> Hankel GEMM (FP64 MFMA) with the
  propagator phase fused in its epilogue -> Hankel GEMM -> IFFT_t. The FFT along t
  commutes with the Hankel matrix along r, so this is the reference's
  ``Y . ifft(phase . fft(Y . u))`` with both GEMMs in the spectral domain.
* ``NonlinearOperator.split(u, rho, rho_s)``: one fused row kernel (pointwise
  Keldysh / plasma terms, Raman convolution and the three spectrally weighted
  transforms in LDS).
* ``NonlinearOperator.update_plasma(...)``: the carrier rate equations as a
  parallel multiple-shooting scan over t (FP64).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch
from scipy.constants import c, epsilon_0

from . import _lib
from ._lib import call, ptr, stream_ptr

SUPPORTED_NT = (512, 1024, 2048, 4096)


def _dev(a, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device)


def _c128(a, device):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    return torch.from_numpy(a).to(device)


def raman_recursion_table(cfg, g):
    """Nt + 2 complex entries: w^m for m = 0..Nt, then (C, 0), where the sampled Raman
    response of grids.py is h[k] = C Im(w^k) (k = 1..Nt/2-1), w = exp(dt (-1/tau_d +
    i/tau_s)), C = dt Omega_r2 tau_s: the kernels evaluate the Raman convolution
    Re(IFFT(FFT(I) R_f)) as the equivalent first-order recursion (csrc/raman.hpp)."""
    Nt, dt = cfg.Nt, g["dt"]
    m = np.arange(Nt + 1)
    out = np.empty(Nt + 2, dtype=np.complex128)
    out[:Nt + 1] = np.exp(m * dt * (-1.0 / cfg.tau_d + 1j / cfg.tau_s))
    out[Nt + 1] = dt * cfg.Omega_r2 * cfg.tau_s
    return out


class DeviceTables:
    """Per-run device tables and kernel parameter blocks."""

    def __init__(self, cfg, g, device=None):
        if cfg.Nt not in SUPPORTED_NT:
            raise ValueError(f"Nt={cfg.Nt} not supported (compiled for {SUPPORTED_NT})")
        _lib.load()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        dev = self.device
        self.cfg, self.g = cfg, g
        Nr, Nt = g["Y"].shape[0], cfg.Nt
        self.Nr, self.Nt = Nr, Nt
        self.Yt = _dev(g["Y"].T, torch.float64, dev)
        m = np.arange(Nt)
        self.tw = _c128(np.exp(-2j * np.pi * m / Nt), dev)
        T, iU = g["T_op"], g["inv_U_nl"]
        self.mA = _dev(T * T * iU, torch.float64, dev)
        self.mB = _dev(-T * iU, torch.float64, dev)
        self.mC = _dev(iU, torch.float64, dev)
        self.Rf = _c128(g["R_f"], dev)
        self.rw_host = raman_recursion_table(cfg, g)
        self.rw = _c128(self.rw_host, dev)
        kel = g["keldysh"]
        sp = kel["f_spline"]
        self.n_knots = len(sp.x)
        self.knots = _dev(sp.x, torch.float64, dev)
        self.pcoef = _dev(sp.c, torch.float64, dev)            # (4, n-1) row-major
        self.W_lut = _dev(kel["W_LUT"], torch.float64, dev)
        self.Ws_lut = _dev(kel["W_LUT_s"], torch.float64, dev)
        self.mask_r = _dev(g["mask_r_1d"], torch.float64, dev)
        self.Rk = g["R"] ** 2 / g["j1last"]
        self.iRk = g["j1last"] / g["R"] ** 2
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self._phase = {}

        pp = _lib.PlasmaParams()
        pp.Nr, pp.Nt, pp.dt = Nr, Nt, g["dt"]
        pp.field_to_I = g["invE2"]
        pp.beta_g, pp.beta_s = g["avalanche_coef"], g["avalanche_coef_s"]
        pp.na = cfg.rho_max
        pp.inv_tau_r, pp.inv_tau_ste = g["inv_taur_eff"], g["inv_tau_ste"]
        pp.enable_ste = int(cfg.enable_ste)
        pp.table_size = kel["NLUT"]
        pp.log_min_I, pp.inv_log_step = kel["logImin"], kel["inv_dlog"]
        pp.Imin, pp.Imax = kel["Imin"], kel["Imax"]
        self.plasma_params = pp

        self.tables = _lib.NLTables(ptr(self.tw), ptr(self.mA), ptr(self.mB), ptr(self.mC),
                                    ptr(self.Rf), ptr(self.knots), ptr(self.pcoef), ptr(self.rw))

    def nl_params(self, dz, absorb=None):
        cfg, g = self.cfg, self.g
        p = _lib.NLParams()
        p.Nr, p.Nt, p.dz = self.Nr, self.Nt, float(dz)
        kel = g["keldysh"]
        p.invE2, p.Imin, p.Imax = g["invE2"], kel["Imin"], kel["Imax"]
        p.photo_pref = 1e6 * cfg.Ui / (cfg.n0 * c * epsilon_0)
        p.rho_max = cfg.rho_max
        p.plasma_pref = (g["sigmaomega"] / 2.0) * 100.0
        en_abs = bool(cfg.enable_plasma_absorption) if absorb is None else bool(absorb)
        pc = (1.0 if cfg.enable_plasma_absorption else 0.0) + \
            (1j * cfg.omega0 * cfg.tau_c if cfg.enable_plasma_defocusing else 0.0)
        p.pc_re, p.pc_im = float(np.real(pc)), float(np.imag(pc))
        p.ste_pref = float(g["ste_pref"])
        p.kerr_pref = 3 * cfg.chi3 * cfg.omega0 ** 2 / (8 * g["komega"] * c ** 2)
        p.f_R = cfg.f_R
        p.en_kerr_inst = int(cfg.enable_kerr_instantaneous)
        p.en_kerr_raman = int(cfg.enable_kerr_raman)
        p.en_pi_loss = int(cfg.enable_photoionization_loss)
        p.en_plasma_absorb = int(en_abs)
        p.n_knots = self.n_knots
        return p

    def phase_half(self, dz):
        """exp(i (delta_k - rho^2 invU / (2 k0)) dz/2) * spec_mask on the (Hankel mode,
        frequency) grid (operators.py:50-51), FP64 on the host, cached per dz."""
        key = float(dz)
        if key not in self._phase:
            g = self.g
            rho2 = (g["rholist"] ** 2)[:, None]
            ph = (g["delta_k"][None, :] - rho2 / (2 * g["komega"]) * g["inv_U_op"][None, :]) * dz / 2
            self._phase[key] = _c128(np.exp(1j * ph) * g["spec_mask"][None, :], self.device)
        return self._phase[key]


def _check_field(u, Nr, Nt):
    if u.dtype != torch.complex128 or tuple(u.shape) != (Nr, Nt) or not u.is_cuda:
        raise ValueError(f"expected a complex128 CUDA tensor of shape ({Nr}, {Nt}), got "
                         f"{u.dtype} {tuple(u.shape)} on {u.device}")
    return u.contiguous()


def _check_real(x, Nr, Nt, name):
    if x.dtype != torch.float64 or tuple(x.shape) != (Nr, Nt) or not x.is_cuda:
        raise ValueError(f"{name}: expected a float64 CUDA tensor of shape ({Nr}, {Nt})")
    if not x.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    return x


class LinearOperator:
    """Diffraction (with space-time focusing) and dispersion over dz/2, diagonal in
    the (Hankel mode, frequency) basis."""

    def __init__(self, cfg, g, tables: DeviceTables = None):
        self.t = tables if tables is not None else DeviceTables(cfg, g)
        self.komega, self.R, self.j1last = g["komega"], g["R"], g["j1last"]
        self.Rk, self.iRk = self.t.Rk, self.t.iRk
        self._work = None

    def half_linear(self, u, dz):
        t = self.t
        u = _check_field(u, t.Nr, t.Nt)
        if self._work is None:
            self._work = torch.empty(2 * t.Nr * t.Nt * 2, dtype=torch.float64, device=t.device)
        out = torch.empty_like(u)
        call("fp_half_linear", ptr(u), ptr(out), t.Nr, t.Nt, ptr(t.Yt), ptr(t.phase_half(dz)),
             t.Rk, t.iRk, ptr(t.tw), ptr(self._work), stream_ptr())
        return out


class NonlinearOperator:
    """Right-hand side of the nonlinear field equation (six switchable channels)
    and the carrier rate equations."""

    def __init__(self, cfg, g, tables: DeviceTables = None):
        self.t = tables if tables is not None else DeviceTables(cfg, g)
        self.cfg = cfg
        self.n0, self.Ui, self.f_R = cfg.n0, cfg.Ui, cfg.f_R
        self.invE2 = g["invE2"]
        self.kerr_pref = 3 * cfg.chi3 * cfg.omega0 ** 2 / (8 * g["komega"] * c ** 2)
        self.plasma_pref = (g["sigmaomega"] / 2.0) * 100.0
        self.plasma_phase = 1.0 + 1j * cfg.omega0 * cfg.tau_c
        self.ste_pref = g["ste_pref"]
        self.avalanche, self.avalanche_s = g["avalanche_coef"], g["avalanche_coef_s"]
        self.inv_taur, self.inv_tau_ste = g["inv_taur_eff"], g["inv_tau_ste"]
        self.rho_max = cfg.rho_max
        self.enable_ste = int(cfg.enable_ste)
        self._params = {}

    def params(self, dz=0.0, absorb=None):
        key = (float(dz), absorb)
        if key not in self._params:
            self._params[key] = self.t.nl_params(dz, absorb)
        return self._params[key]

    def split(self, u, rho, rho_s=None):
        t = self.t
        u = _check_field(u, t.Nr, t.Nt)
        _check_real(rho, t.Nr, t.Nt, "rho")
        if rho_s is not None:
            _check_real(rho_s, t.Nr, t.Nt, "rho_s")
        rhs = torch.empty_like(u)
        alpha = torch.empty(t.Nr, t.Nt, dtype=torch.float64, device=t.device)
        call("fp_nl_split", ptr(u), ptr(rho), ptr(rho_s), ptr(rhs), ptr(alpha),
             ctypes.byref(self.params()), ctypes.byref(t.tables), stream_ptr())
        return rhs, alpha

    def loss_rates(self, u, rho):
        """(photo-ionisation, plasma-absorption) field-loss rates in 1/m
        (operators.py:148-163): photo is split()'s alpha with the plasma term off."""
        t = self.t
        u = _check_field(u, t.Nr, t.Nt)
        _check_real(rho, t.Nr, t.Nt, "rho")
        rhs = torch.empty_like(u)
        photo = torch.empty(t.Nr, t.Nt, dtype=torch.float64, device=t.device)
        call("fp_nl_split", ptr(u), ptr(rho), None, ptr(rhs), ptr(photo),
             ctypes.byref(self.params(0.0, absorb=False)), ctypes.byref(t.tables), stream_ptr())
        plasma = self.plasma_pref * rho if self.cfg.enable_plasma_absorption else torch.zeros_like(photo)
        return photo, plasma

    def update_plasma(self, u, rho, rho_s, dt=None, blocks=None, threads=None):
        """Rate equations along t from the current field; overwrites rho, rho_s
        (column 0 = 0). dt/blocks/threads are accepted for signature
        compatibility; the grid's dt and the kernel's own launch shape are used."""
        t = self.t
        u = _check_field(u, t.Nr, t.Nt)
        _check_real(rho, t.Nr, t.Nt, "rho")
        _check_real(rho_s, t.Nr, t.Nt, "rho_s")
        pp = t.plasma_params
        if dt is not None and abs(dt - pp.dt) > 1e-12 * abs(pp.dt):
            pp = _lib.PlasmaParams.from_buffer_copy(pp)
            pp.dt = float(dt)
        call("fp_plasma_scan", ptr(u), ptr(rho), ptr(rho_s), ctypes.byref(pp), ptr(t.W_lut),
             ptr(t.Ws_lut), ptr(t.status), stream_ptr())
