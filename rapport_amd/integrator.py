"""Z-marching driver: drop-in for ``Integrator`` of ``D3/sim/integrator.py:28-340``
(D3 = ``.github/deliverable (3)``).

Same loop semantics as the reference: ``propagate()`` runs ``nz + 1`` iterations;
iteration i recomputes the carrier densities from the current field (frozen
for the whole step), applies the Strang step
``L/2 -> alpha/2 -> RK4(N) -> alpha/2 -> L/2 -> x mask_r`` and records the state
with the label ``z = begin + i dz``. Outputs (``result.npz`` keys and
``params.json``) follow integrator.py:238-340.

GPU pipeline of one iteration (all on one stream, no host synchronisation):

    plasma scan   spectral field -> IFFT_t (LDS) -> rho, rho_s          [row kernel]
    L/2           G1: Rk Y @ . x phase   G2: iRk Y @ .                   [FP64 MFMA GEMMs]
    N             IFFT_t -> alpha/2 -> RK4 (4 x split) -> alpha/2 -> FFT_t [row kernel]
    L/2 + mask    G3: Rk Y @ . x phase   G4: iRk mask_r Y @ .

The field is kept between iterations as its FFT along t (the Hankel matrix acts
along r and commutes with the FFT along t, and FFT(IFFT(x)) = x), so the only
time-domain copies formed are inside the row kernels; ``Integrator.u`` converts
on demand. Recording, when due, adds one inverse FFT and one reduction kernel
and writes into device-resident histories.
"""
from __future__ import annotations

import ctypes
import json
import os
import weakref
from typing import Callable, Union

import numpy as np
import torch

from . import _lib
from ._lib import RecordDst, call, ptr, stream_ptr
from .config import Config, code_fingerprint
from .grids import ENVELOPES, build_grids
from .operators import DeviceTables, LinearOperator, NonlinearOperator


class Integrator:
    def __init__(self, cfg: Config, envelope: Union[str, Callable] = "gaussian_focused",
                 device=None):
        self.cfg = cfg
        if cfg.write_output:
            os.makedirs(cfg.out_dir, exist_ok=True)
        g = build_grids(cfg)
        self.g = g
        self.dz, self.dt = cfg.dz, g["dt"]
        self.t = DeviceTables(cfg, g, device)
        dev = self.t.device
        self.device = dev
        self.lin = LinearOperator(cfg, g, self.t)
        self.nl = NonlinearOperator(cfg, g, self.t)

        fn = envelope if callable(envelope) else ENVELOPES[envelope]
        u0 = fn(g["rr"], g["tt"], cfg, g)
        if isinstance(u0, torch.Tensor):
            u0 = u0.detach().cpu().numpy()
        u0 = np.ascontiguousarray(np.asarray(u0).astype(np.complex128, copy=False))
        self.Nr, self.Nt = u0.shape
        u0d = torch.from_numpy(u0).to(dev)
        from .mixed import zero_columns
        self._zc = zero_columns(g["spec_mask"])   # real columns where the mask is exactly 0
        self._spec = None                       # FFT_t of the field (persistent buffer)
        self._u_last = None                     # (weakref, version) of the last `u` handed out
        self.rho = torch.zeros(self.Nr, self.Nt, dtype=torch.float64, device=dev)
        self.rho_s = torch.zeros(self.Nr, self.Nt, dtype=torch.float64, device=dev)
        self.mask_r = self.t.mask_r
        self._mx = None
        if cfg.precision == "mixed":
            from .mixed import MixedEngine
            self._mx = MixedEngine(self.t, cfg, g, u0d)
        else:
            self._spec = torch.empty_like(u0d)
            call("fp_fft_rows", ptr(u0d), ptr(self._spec), self.Nr, self.Nt, 0, ptr(self.t.tw),
                 stream_ptr())
        del u0d

        r = g["rlist"]
        self._r_w = 2.0 * np.pi * r * np.diff(r, prepend=0.0)   # 2 pi r dr (rectangle)
        fl0 = np.sum(np.abs(u0) ** 2, axis=1) * self.dt * g["invE2"] * 1e4
        self.U0_uJ = float(np.sum(fl0 * self._r_w)) * 1e6
        if cfg.verbose:
            print(f"[init] U_beam(0) = {self.U0_uJ:.3f} uJ ", flush=True)

        n_saves = cfg.nz // cfg.save_stride + 1
        self.n_saves = n_saves
        f32 = dict(dtype=torch.float32, device=dev)
        self.fluence_rz = torch.zeros(n_saves, self.Nr, **f32)
        self.rho_rz = torch.zeros(n_saves, self.Nr, **f32)
        self.rho_s_rz = torch.zeros(n_saves, self.Nr, **f32)
        self.Imax_z = np.zeros(n_saves)
        self._Imax_dev = torch.zeros(n_saves, dtype=torch.float64, device=dev)
        self.z_saved = np.zeros(n_saves)
        self._dEdz_dev = torch.zeros(2, n_saves, dtype=torch.float64, device=dev)
        self._dEdz_photo_uJm = np.zeros(n_saves)
        self._dEdz_plasma_uJm = np.zeros(n_saves)
        self.E_STE_z = np.zeros(n_saves)
        self.k_save = 0
        self.t_full_fs = g["tlist"] * 1e15

        self._t_stride = max(1, cfg.rho_t_stride) if cfg.rho_t_stride > 0 else 0
        self._r_stride = max(1, cfg.rho_r_stride)
        if self._t_stride > 0:
            Nt_sub = (self.Nt - 1) // self._t_stride + 1
            self._Nt_sub = Nt_sub
            Nr_sub = (self.Nr - 1) // self._r_stride + 1
            self.rho_rzt = torch.zeros(n_saves, Nr_sub, Nt_sub, **f32)
            self.rho_s_rzt = torch.zeros(n_saves, Nr_sub, Nt_sub, **f32)
            self.I_rzt = torch.zeros(n_saves, Nr_sub, Nt_sub, **f32)
            self.t_sub_fs = self.t_full_fs[::self._t_stride]
            self.r_sub = r[::self._r_stride]
        else:
            self.rho_rzt = self.rho_s_rzt = self.I_rzt = None
            self.t_sub_fs = self.r_sub = None
        cube_b = 3 * 4 * self.rho_rzt.numel() if self.rho_rzt is not None else 0
        rest_b = 4 * n_saves * (10 * 2 * self.Nr + 3 * self.Nt)
        self.size_estimate_gb = (cube_b + rest_b) / 1024 ** 3
        if cfg.verbose:
            print(f"[init] result.npz ~ {self.size_estimate_gb:.2f} GB non compresse"
                  + (f" ({100 * cube_b / (cube_b + rest_b):.0f} % = cube (z,r,t))" if cube_b else ""),
                  flush=True)
        if self.size_estimate_gb > 2.0:
            print(f"[WARN] result.npz estimated at {self.size_estimate_gb:.1f} GB; the (z,r,t) cube is "
                  f"set by rho_t_stride={cfg.rho_t_stride} / rho_r_stride={cfg.rho_r_stride}; "
                  "rho_t_stride=0 disables it.", flush=True)

        self.absorbed_rz = torch.zeros(n_saves, self.Nr, **f32)
        tl = self.t_full_fs
        if cfg.absorb_time_bins_fs:
            edges = np.asarray(cfg.absorb_time_bins_fs, dtype=np.float64)
            lo, hi = [], []
            for a, b in zip(edges[:-1], edges[1:]):
                idx = np.nonzero((tl >= a) & (tl < b))[0]   # tlist increases: a contiguous range
                lo.append(int(idx[0]) if idx.size else 0)
                hi.append(int(idx[-1]) + 1 if idx.size else 0)
            self._bin_lo = torch.tensor(lo, dtype=torch.int32, device=dev)
            self._bin_hi = torch.tensor(hi, dtype=torch.int32, device=dev)
            self.absorb_bin_edges_fs = edges
            self.absorbed_rz_bins = torch.zeros(n_saves, len(lo), self.Nr, **f32)
        else:
            self._bin_lo = self._bin_hi = None
            self.absorb_bin_edges_fs = self.absorbed_rz_bins = None
        self.Ipeak_rz = torch.zeros(n_saves, self.Nr, **f32)
        if cfg.rho_snapshot_t_fs is not None:
            self._it_snap = int(np.argmin(np.abs(tl - cfg.rho_snapshot_t_fs)))
            self.rho_rz_at = torch.zeros(n_saves, self.Nr, **f32)
        else:
            self._it_snap, self.rho_rz_at = None, None
        self.rho_onaxis_t = torch.zeros(n_saves, self.Nt, **f32)
        self.rho_s_onaxis_t = torch.zeros(n_saves, self.Nt, **f32)
        self.I_onaxis_t = torch.zeros(n_saves, self.Nt, **f32)

        self._rec_partial = torch.zeros(3, self.Nr, dtype=torch.float64, device=dev)
        self._r_w_dev = torch.from_numpy(self._r_w).to(dev)
        self._w0 = self._w1 = None
        self._utmp = None

    # ------------------------------------------------------------------ state
    # The field lives in one persistent buffer as its FFT along t (complex128, or
    # complex64 in the mixed engine); kernels captured in a graph keep pointing at it.
    @property
    def u(self):
        """Time-domain field (Nr x Nt complex128 on the GPU; the mixed path upcasts
        its complex64 state), a new tensor each call. As with the reference's array
        (integrator.py:40), writing into it in place (``it.u[...] = x``) sets the field
        that the next step, record or propagate() starts from."""
        self._sync_user_edit()
        if self._mx is not None:
            out = self._mx.time_c64().to(torch.complex128)
        else:
            out = torch.empty_like(self._spec)
            call("fp_fft_rows", ptr(self._spec), ptr(out), self.Nr, self.Nt, 1, ptr(self.t.tw),
                 stream_ptr())
        self._u_last = (weakref.ref(out), out._version)
        return out

    @u.setter
    def u(self, value):
        self._u_last = None
        if self._mx is not None:
            self._mx.set_time(value)
            return
        v = value.to(torch.complex128).contiguous()
        call("fp_fft_rows", ptr(v), ptr(self._spec), self.Nr, self.Nt, 0, ptr(self.t.tw),
             stream_ptr())

    def _sync_user_edit(self):
        """Push an in-place edit of the last tensor ``u`` returned into the state."""
        last = self._u_last
        if last is None:
            return
        t = last[0]()
        if t is None:
            self._u_last = None
        elif t._version != last[1]:
            self.u = t

    def _spectral(self):
        return self._spec

    def _rho_s_arg(self):
        return self.rho_s if (self.cfg.enable_ste and self.g["ste_pref"]) else None

    # ------------------------------------------------------------------ one step
    def step(self, dz):
        """Integrator.step (integrator.py:138-149) on the time-domain field."""
        self._sync_user_edit()
        if self._mx is not None:
            self._mx.strang(dz, self.rho, self._rho_s_arg(), self.nl.params(dz))
            return
        u = self.lin.half_linear(self.u, dz)
        call("fp_nl_step", ptr(u), ptr(u), ptr(self.rho), ptr(self._rho_s_arg()), 0, 0,
             ctypes.byref(self.nl.params(dz)), ctypes.byref(self.t.tables), stream_ptr())
        u = self.lin.half_linear(u, dz)
        self.u = u * self.mask_r[:, None]

    def _pipeline_step(self, dz):
        """One propagate() iteration body (plasma + step) on the spectral state."""
        self._sync_user_edit()
        if self._mx is not None:
            self._mx.step(dz, self.rho, self.rho_s, self._rho_s_arg(), self.nl.params(dz))
            return
        t = self.t
        spec = self._spectral()
        if self._w0 is None:
            self._w0 = torch.empty_like(spec)
            self._w1 = torch.empty_like(spec)
        w0, w1 = self._w0, self._w1
        st = stream_ptr()
        pp = t.plasma_params
        rho_s = self.rho_s if self.cfg.enable_ste else None   # stays zero without STE
        call("fp_plasma_scan_spec", ptr(spec), ptr(self.rho), ptr(rho_s), None,
             ctypes.byref(pp), ptr(t.W_lut), ptr(t.Ws_lut), ptr(t.tw), ptr(t.status), st)
        phase = t.phase_half(dz)
        # GEMMs only over the frequencies where the spectral mask is not 0 (the product is
        # exactly 0 elsewhere); G2 / G4 write those zeros, as the row kernels read whole rows
        zlo, zhi = self._zc
        Nr, Nt = self.Nr, self.Nt
        call("fp_hankel_gemm_band", ptr(t.Yt), ptr(spec), ptr(w1), Nr, Nt, t.Rk, None,
             ptr(phase), zlo, zhi, 0, st)
        call("fp_hankel_gemm_band", ptr(t.Yt), ptr(w1), ptr(w0), Nr, Nt, t.iRk, None, None,
             zlo, zhi, 1, st)
        call("fp_nl_step", ptr(w0), ptr(w0), ptr(self.rho), ptr(self._rho_s_arg()), 1, 1,
             ctypes.byref(self.nl.params(dz)), ctypes.byref(t.tables), st)
        call("fp_hankel_gemm_band", ptr(t.Yt), ptr(w0), ptr(w1), Nr, Nt, t.Rk, None,
             ptr(phase), zlo, zhi, 0, st)
        call("fp_hankel_gemm_band", ptr(t.Yt), ptr(w1), ptr(spec), Nr, Nt, t.iRk,
             ptr(t.mask_r), None, zlo, zhi, 1, st)

    def capture_step(self, steps=1):
        """Capture `steps` z-steps (plasma scan + Strang step, as propagate() runs them)
        into a HIP graph on a side stream and return it; graph.replay() then advances
        the persistent state by that many steps. The capture itself does not advance
        the state (one eager warm-up step runs first and allocates the work buffers)."""
        self._sync_user_edit()
        snap = self._snapshot()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._pipeline_step(self.dz)
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=side):
            for _ in range(steps):
                self._pipeline_step(self.dz)
        torch.cuda.synchronize(self.device)
        self._restore(snap)
        return graph

    def _snapshot(self):
        st = (self._mx.spec if self._mx is not None else self._spec).clone()
        return st, self.rho.clone(), self.rho_s.clone()

    def _restore(self, snap):
        st, r, rs = snap
        (self._mx.spec if self._mx is not None else self._spec).copy_(st)
        self.rho.copy_(r)
        self.rho_s.copy_(rs)
        if self._mx is not None:
            self._mx.slots.zero_()
            call("fp_absmax_f32", ptr(self._mx.spec), 2 * self.Nr * self.Nt, self._mx._slot(0),
                 stream_ptr())

    def advance(self, i):
        """Body of the propagate() loop for iteration i (integrator.py:157-162)."""
        self._pipeline_step(self.dz)
        if i % self.cfg.save_stride == 0:
            self._record(self.cfg.begin + i * self.dz)

    def propagate(self, n_steps=None, graph=False):
        """The z-loop (integrator.py:151-177): nz + 1 iterations (or n_steps). With
        graph=True each iteration is one replay of a captured HIP graph of the step."""
        cfg = self.cfg
        n = cfg.nz + 1 if n_steps is None else int(n_steps)
        g1 = self.capture_step(1) if graph else None
        it = range(n)
        pbar = None
        if cfg.verbose:
            try:
                from tqdm.auto import tqdm
                pbar = tqdm(it, desc="Filamentation", unit="step")
                it = pbar
            except ImportError:
                pass
        for i in it:
            if g1 is not None:
                g1.replay()
                if i % cfg.save_stride == 0:
                    self._record(cfg.begin + i * self.dz)
            else:
                self.advance(i)
            if pbar is not None and i > 0 and i % cfg.ckpt_every == 0 and self.k_save > 0:
                k = self.k_save - 1
                flu = self.fluence_rz[k].double().cpu().numpy()
                U = float(np.sum(flu * self._r_w)) * 100.0
                pbar.set_postfix(z=f"{(cfg.begin + i * self.dz) * 1e6:+.0f}um",
                                 U=f"{U / self.U0_uJ * 100:.1f}%",
                                 I_peak=f"{float(self._Imax_dev[k]):.2e}")
        if pbar is not None:
            pbar.close()
        if int(self.t.status.item()) != 0:
            print(f"[WARN] plasma scan hit its iteration cap on {int(self.t.status.item())} rows")
        return self._results()

    # ------------------------------------------------------------------ recording
    def _record(self, z_now):
        """integrator.py:179-224 as one fp_record_history call: the inverse transform of
        the spectral state, the reductions of loss_rates (operators.py:148-163) and every
        history entry of record k are written on the device."""
        self._sync_user_edit()
        k = self.k_save
        if not 0 <= k < self.n_saves:
            raise IndexError(f"record {k}: the histories hold {self.n_saves} records")
        if self._mx is not None:
            spec, c64, tw = self._mx.spec, 1, self._mx.m.tw
        else:
            spec, c64, tw = self._spec, 0, self.t.tw
        nb = 0 if self._bin_lo is None else int(self._bin_lo.numel())
        call("fp_record_history", ptr(spec), c64, ptr(tw), ptr(self.rho), ptr(self.rho_s),
             ctypes.byref(self.nl.params(self.dz)), ctypes.byref(self.t.tables), self.dt, nb,
             ptr(self._bin_lo), ptr(self._bin_hi), ctypes.byref(self._record_dst(k)),
             stream_ptr())
        self.z_saved[k] = float(z_now)
        self.k_save += 1

    def _record_dst(self, k):
        """Device addresses of the entries at index k of every history."""
        def at(x):
            return None if x is None else x.data_ptr() + k * x.stride(0) * x.element_size()
        cube = self._t_stride > 0 and self.rho_rzt is not None
        dE = self._dEdz_dev.data_ptr()
        return RecordDst(
            fluence=at(self.fluence_rz), rho=at(self.rho_rz), rho_s=at(self.rho_s_rz),
            ipeak=at(self.Ipeak_rz), absorbed=at(self.absorbed_rz), bins=at(self.absorbed_rz_bins),
            rho_at=at(self.rho_rz_at), it_snap=self._it_snap or 0,
            rho_onaxis=at(self.rho_onaxis_t), rho_s_onaxis=at(self.rho_s_onaxis_t),
            I_onaxis=at(self.I_onaxis_t), r_stride=self._r_stride,
            t_stride=self._t_stride if cube else 0, Nt_sub=self._Nt_sub if cube else 0,
            rho_cube=at(self.rho_rzt), rho_s_cube=at(self.rho_s_rzt), I_cube=at(self.I_rzt),
            r_w=self._r_w_dev.data_ptr(), partial=self._rec_partial.data_ptr(),
            imax=at(self._Imax_dev), dedz_photo=dE + 8 * k, dedz_plasma=dE + 8 * (self.n_saves + k))

    def _cumulative_energy_fraction(self, dEdz_uJm):
        """Trapezoid cumulative integral of dE/dz over z, over U0 (integrator.py:226-236)."""
        z = self.z_saved[:self.k_save]
        d = dEdz_uJm[:self.k_save]
        if len(z) < 2:
            return np.zeros_like(d)
        seg = 0.5 * (d[:-1] + d[1:]) * np.diff(z)
        return np.concatenate([[0.0], np.cumsum(seg)]) / self.U0_uJ

    def _results(self):
        """integrator.py:238-298: mirrored radial axis, cumulative losses, atomic npz."""
        cfg, g, k = self.cfg, self.g, self.k_save
        self.Imax_z[:] = self._Imax_dev.cpu().numpy()
        dE = self._dEdz_dev.cpu().numpy()
        self._dEdz_photo_uJm[:], self._dEdz_plasma_uJm[:] = dE[0], dE[1]
        host = lambda x: x[:k].cpu().numpy() if x is not None else None  # noqa: E731
        mirror = lambda a: np.hstack([a[:, ::-1], a])                     # noqa: E731
        r = g["rlist"]
        E_MPI = self._cumulative_energy_fraction(self._dEdz_photo_uJm)
        E_pl = self._cumulative_energy_fraction(self._dEdz_plasma_uJm)
        flu, rr, rs = host(self.fluence_rz), host(self.rho_rz), host(self.rho_s_rz)
        out = dict(
            r=np.concatenate([-r[::-1], r]), rlist=r, z=self.z_saved[:k],
            fluence_rz=mirror(flu), rho_rz=mirror(rr), rho_s_rz=mirror(rs),
            Imax_z=self.Imax_z[:k], E_plasma_z=E_pl, E_MPI_z=E_MPI, E_STE_z=self.E_STE_z[:k],
            E_total_z=E_MPI + E_pl,
            rho_rzt=host(self.rho_rzt), rho_s_rzt=host(self.rho_s_rzt), I_rzt=host(self.I_rzt),
            t_sub_fs=self.t_sub_fs, r_sub=self.r_sub,
            absorbed_rz=mirror(host(self.absorbed_rz)), Ipeak_rz=mirror(host(self.Ipeak_rz)),
            rho_rz_at=mirror(host(self.rho_rz_at)) if self.rho_rz_at is not None else None,
            absorbed_rz_bins=(np.stack([mirror(b) for b in
                                        np.moveaxis(host(self.absorbed_rz_bins), 1, 0)], axis=1)
                              if self.absorbed_rz_bins is not None else None),
            absorb_bin_edges_fs=self.absorb_bin_edges_fs,
            rho_onaxis_t=host(self.rho_onaxis_t), rho_s_onaxis_t=host(self.rho_s_onaxis_t),
            I_onaxis_t=host(self.I_onaxis_t), t_full_fs=self.t_full_fs,
        )
        if cfg.write_output:
            final = os.path.join(cfg.out_dir, "result.npz")
            tmp = final + ".tmp.npz"
            try:
                np.savez_compressed(tmp, **out)
                os.replace(tmp, final)
            except BaseException:
                for f in (tmp, tmp + ".npz"):
                    try:
                        os.remove(f)
                    except OSError:
                        pass
                raise
            self._dump_params()
        return out

    def _dump_params(self):
        """params.json (integrator.py:300-340)."""
        cfg = self.cfg
        params = dict(
            n0=cfg.n0, n0_probe=cfg.n0_probe, nc_probe_cm3=cfg.nc_probe,
            lambda_probe_nm=cfg.lambda_probe * 1e9, wavelength_nm=cfg.wavelength * 1e9,
            n2=cfg.n2, U_g_eV=cfg.Ui_eV, Us_eV=cfg.Us_eV, energy_uJ=cfg.energy_uJ,
            w0_um=cfg.w0 * 1e6, delta_t_fs=cfg.delta_t * 1e15,
            code_fingerprint=code_fingerprint(),
            begin_um=cfg.begin * 1e6, end_um=cfg.end * 1e6, z_focus_air_um=cfg.z_focus_air_um,
            toggles=dict(
                enable_kerr_instantaneous=cfg.enable_kerr_instantaneous,
                enable_kerr_raman=cfg.enable_kerr_raman,
                enable_self_steepening=cfg.enable_self_steepening,
                enable_photoionization_loss=cfg.enable_photoionization_loss,
                enable_plasma_defocusing=cfg.enable_plasma_defocusing,
                enable_plasma_absorption=cfg.enable_plasma_absorption,
                enable_space_time_focusing=cfg.enable_space_time_focusing,
                enable_spectral_filter=cfg.enable_spectral_filter,
                tau_ste_fs=(cfg.tau_ste * 1e15 if cfg.tau_ste else None),
                enable_ste_index=cfg.enable_ste_index,
                E_tr_eV=cfg.E_tr_eV, f_ste_pump=float(self.g["f_ste"]),
                enable_avalanche=cfg.enable_avalanche,
                enable_recombination=cfg.enable_recombination,
                enable_ste=cfg.enable_ste,
            ),
        )
        with open(os.path.join(cfg.out_dir, "params.json"), "w") as f:
            json.dump(params, f, indent=2)
