"""Batched form: many independent mixed-precision propagations advanced together.

Each instance is an ordinary mixed-precision ``Integrator`` (its own Config, grid
shape, pulse, tables, spectral state and histories). ``BatchedIntegrator``
describes all of them to the library once (``fp_batch_plan_create``) and then
advances every instance by one z-step with six grouped passes
(plasma scan, G1, G2, nonlinear step, G3, G4): each pass is one launch per group
of instances that share a kernel instantiation (row kernels: same Nt and STE flag;
GEMM: same tile shape), whose blocks map to (instance, row / tile) through a
prefix table, optionally as a persistent grid. The whole z-step is captured in
one HIP graph and replayed. Per-instance arithmetic is that of the single-instance
mixed sequence (the same device functions), so each instance's result equals its
own single run.

Recording (``save_stride``) uses each member's own reduction kernel when due.
"""
from __future__ import annotations

import ctypes
from dataclasses import replace
from typing import Callable, List, Sequence, Union

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_ptr
from .config import Config
from .integrator import Integrator
from .mixed import zero_columns  # noqa: F401  (re-exported: the batch descriptors use it)

PASSES = dict(plasma=0, G1=1, G2=2, nl=3, G3=4, G4=5)
KIND = dict(plasma=0, gemm=1, nl=2)


def _vp(t):
    return None if t is None else t.data_ptr()


class BatchedIntegrator:
    """``members``: list of mixed-precision Integrators advanced in lock-step.

    Parameters
    ----------
    cfgs: Configs of the instances (precision is forced to "mixed").
    envelopes: one envelope (name or callable) for all, or one per instance.
    tune: run the autotuner (``rapport_amd.autotune.tune``) on the live batch.
    """

    def __init__(self, cfgs: Sequence[Config],
                 envelopes: Union[str, Callable, Sequence] = "gaussian_focused", device=None,
                 tune: bool = False):
        n = len(cfgs)
        if n == 0:
            raise ValueError("empty batch")
        if isinstance(envelopes, (str,)) or callable(envelopes):
            envelopes = [envelopes] * n
        if len(envelopes) != n:
            raise ValueError("one envelope per instance expected")
        self.members: List[Integrator] = []
        for cfg, env in zip(cfgs, envelopes):
            c = cfg if cfg.precision == "mixed" else replace(cfg, precision="mixed")
            self.members.append(Integrator(c, envelope=env, device=device))
        self.device = self.members[0].device
        self.n = n
        self._descs = (_lib.BatchInstance * n)()
        self._keep = []
        for i, m in enumerate(self.members):
            self._fill(self._descs[i], m)
        self._plan = ctypes.c_void_p()
        call("fp_batch_plan_create", self._descs, n, ctypes.byref(self._plan))
        self.config = {}
        self._graph = None
        if tune:
            from .autotune import tune as _tune
            _tune(self)

    # ------------------------------------------------------------------ plan
    def _fill(self, d, m: Integrator):
        mx, t = m._mx, m.t
        mt = mx.m
        if mx.Np != m.Nr:
            raise ValueError(f"the batched form needs Nr % 128 == 0 (an instance has Nr = {m.Nr})")
        phase = mt.phase_half(m.dz)
        nlp = m.nl.params(m.dz)
        self._keep += [phase, nlp]
        ste = bool(m.cfg.enable_ste)
        d.Nr, d.Nt = m.Nr, m.Nt
        d.y_scale_exp = mt.y_scale_exp
        d.enable_ste = int(ste)
        d.Rk, d.iRk = mt.Rk, mt.iRk
        d.Yh, d.Yl = _vp(mt.Yh), _vp(mt.Yl)
        d.spec, d.w0, d.w1 = _vp(mx.spec), _vp(mx.w0), _vp(mx.w1)
        d.rho, d.rho_s = _vp(m.rho), _vp(m.rho_s)
        d.phase, d.mask_r, d.slots = _vp(phase), _vp(mt.mask_r), _vp(mx.slots)
        d.W_lut, d.Ws_lut, d.tw, d.status = _vp(t.W_lut), _vp(t.Ws_lut), _vp(mt.tw), _vp(t.status)
        d.pp = t.plasma_params
        d.np = nlp
        d.nt = mt.tables
        d.zc_lo, d.zc_hi = zero_columns(m.g["spec_mask"])

    def set_config(self, i, bm=0, bn=0, nl_tpb=0):
        """Launch configuration of instance i (GEMM tile bm x bn, nonlinear threads)."""
        call("fp_batch_set_config", self._plan, int(i), int(bm), int(bn), int(nl_tpb))
        self._graph = None

    def set_grid(self, kind, blocks):
        """Grid of pass kind ('plasma', 'gemm', 'nl'): 0 = one block per work item,
        > 0 = that many persistent blocks."""
        call("fp_batch_set_grid", self._plan, KIND[kind], int(blocks))
        self._graph = None

    def set_merged(self, merged=True):
        """One launch per pass for the whole batch (six launches per z-step) instead of
        one per group of instances sharing a kernel configuration."""
        call("fp_batch_set_merged", self._plan, int(bool(merged)))
        self.merged = 0x3F if merged else 0
        self._graph = None

    def set_merged_passes(self, mask):
        """Launch mode per pass: bit p (pass p: 0 scan, 1 G1, 2 G2, 3 nonlinear step, 4 G3,
        5 G4) set = one launch for the whole batch, clear = grouped launches."""
        call("fp_batch_set_merged_passes", self._plan, int(mask))
        self.merged = int(mask)
        self._graph = None

    def launches_per_step(self):
        return int(_lib.load().fp_batch_launches(self._plan))

    def close(self):
        if self._plan:
            _lib.load().fp_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:   # interpreter shutdown
            pass

    # ------------------------------------------------------------------ stepping
    def run_pass(self, name):
        call("fp_batch_pass", self._plan, PASSES[name], stream_ptr())

    def step(self):
        """One z-step (plasma scan + Strang step) of every instance, eagerly."""
        call("fp_batch_step_mixed", self._plan, stream_ptr())

    def _snapshot(self):
        """State of every instance: the spectral field and the carrier densities (the
        previous densities seed the carrier scan's Newton iteration)."""
        return [(m._mx.spec.clone(), m.rho.clone(), m.rho_s.clone()) for m in self.members]

    def _restore(self, snap):
        for m, (s, r, rs) in zip(self.members, snap):
            m._mx.spec.copy_(s)
            m.rho.copy_(r)
            m.rho_s.copy_(rs)
        self.refresh_scales()

    def refresh_scales(self):
        """Recompute each instance's slot 0 (max|spec|) after touching the state."""
        for m in self.members:
            m._mx.slots.zero_()
            call("fp_absmax_f32", ptr(m._mx.spec), 2 * m.Nr * m.Nt, m._mx._slot(0), stream_ptr())

    def capture(self, steps=1):
        """HIP graph of `steps` batched z-steps; capturing does not advance the state
        (one eager warm-up step runs first and is undone)."""
        snap = self._snapshot()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.step()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=side):
            for _ in range(steps):
                self.step()
        torch.cuda.synchronize(self.device)
        self._restore(snap)
        del snap
        return graph

    def propagate(self, n_steps=None, graph=True):
        """Run every instance for n_steps z-steps (default: the largest nz + 1),
        recording each one at its own save_stride; returns the members' results."""
        n = max(m.cfg.nz + 1 for m in self.members) if n_steps is None else int(n_steps)
        g = None
        if graph:
            if self._graph is None:
                self._graph = self.capture(1)
            g = self._graph
        for i in range(n):
            if g is not None:
                g.replay()
            else:
                self.step()
            for m in self.members:
                if i % m.cfg.save_stride == 0 and m.k_save < m.n_saves:
                    m._record(m.cfg.begin + i * m.dz)
        return [m._results() for m in self.members]

    # ------------------------------------------------------------------ bookkeeping
    def points(self):
        """Total grid points (sum of Nr x Nt)."""
        return sum(m.Nr * m.Nt for m in self.members)

    def gemm_flops(self):
        """Algorithmic flops of the four Hankel GEMMs of one batched step."""
        return sum(16.0 * m.Nr * m.Nr * m.Nt for m in self.members)


def random_sweep(n=64, seed=0, nz=1999, dz=10e-6, Nr_choices=(256, 512, 768, 1024),
                 Nt_choices=(1024, 2048, 4096), energy_uJ=(0.5, 20.0), fwhm_fs=(40.0, 200.0),
                 w0_um=(50.0, 150.0), noise=0.01, **cfg_kw):
    """Configs and envelopes of a sweep of n instances: Nr and Nt drawn from the
    choice lists, energy log-uniform, FWHM and w0 uniform, each input Gaussian
    (collimated) times (1 + noise * seeded complex Gaussian noise). The time window
    is the reference's 10 tp of each pulse; R = 10 w0."""
    from .grids import envelope_lens_noise
    rng = np.random.default_rng(seed)
    cfgs, envs = [], []
    base = dict(wavelength=800e-9, n2=3.2e-20, Ui_eV=9.0, rho_max=2.1e22, tau_c=1.27e-15,
                tau_r=150e-15, enable_ste=False, R_factor=10.0, begin=0.0, rho_t_stride=0,
                write_output=False, verbose=False, precision="mixed")
    base.update(cfg_kw)
    for k in range(n):
        Nr = int(rng.choice(Nr_choices))
        Nt = int(rng.choice(Nt_choices))
        E = float(np.exp(rng.uniform(np.log(energy_uJ[0]), np.log(energy_uJ[1]))))
        fw = float(rng.uniform(*fwhm_fs)) * 1e-15
        w0 = float(rng.uniform(*w0_um)) * 1e-6
        kw = dict(base, N=Nr + 1, Nt=Nt, nz=nz, end=base["begin"] + nz * dz, energy_uJ=E,
                  delta_t=fw, w0=w0)
        kw.setdefault("save_stride", max(1, nz // 10))
        cfgs.append(Config(**kw))
        envs.append(envelope_lens_noise(lens_f=None, noise=noise, seed=1000 + k))
    return cfgs, envs
