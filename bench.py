"""Benchmark: z-steps/s of the split-step propagator at Nr x Nt = 1024 x 4096,
mixed precision vs FP64, gated on fluence rel-L2 <= 1e-4.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision both|fp64|mixed]
                    [--no-cpu-baseline] [--batched N] [--copy-calibration]

Workload (synthetic input, generated on the host and uploaded before timing):
fused silica at 800 nm, Nr = 1024 radial rows (Hankel order 1025, R = 1.5 mm),
Nt = 4096 samples over a 1.6 ps window, dz = 5 um, 5 uJ / 85 fs / w0 = 100 um
Gaussian with a f = 5 cm lens phase and 1 % seeded complex noise, STE off. A
step is one iteration of the propagate() loop (carrier scan + Strang step) with
recording off; the field stays resident in HBM.

`value` is the mixed-precision z-steps/s when its fluence after `--acc-steps`
steps (recorded every 40 steps) is within rel-L2 1e-4 of the FP64 run of the
same input, otherwise the FP64 z-steps/s; it is timed over K plain steps (no
per-kernel events). `roofline` is for the Hankel GEMM (4 launches per step) of the
headline path: its average launch time is measured over K more steps with HIP
event pairs on the launch stream around every GEMM launch (the event records slow
the step by about 5 %, reported as `ms_per_step_gemm_events`); `roofline_nongemm`
applies the HBM model of the non-GEMM passes (compulsory bytes per point) to the
rest of that step. `cpu_baseline` times the NumPy FP64 oracle (a
restatement of the reference algorithm) on the host for a bounded sample.
`batched` reports the batched form on a 64-instance sweep (Nr in {256, 512,
768, 1024}, Nt in {1024, 2048, 4096}, 0.5-20 uJ, 40-200 fs, w0 50-150 um, 1 %
noise, dz 10 um): instance z-steps/s of the graph-captured batched step, the
time of each pass and the HBM fraction of the non-GEMM passes. `adjoint` times one
gradient (mixed forward with checkpoints + hand-written adjoint) at 512 x 2048 x 1000
steps, checks 4 of its 32 components against FP64 central differences, and runs the
input-profile fit (L-BFGS on the 32 coefficients towards the FP64 target fluence).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 2500.0}     # FP64 MFMA; dense FP16 MFMA
NONGEMM_BYTES_PER_PT = {"fp64": 7 * 16 + 2 * 8, "mixed": 7 * 8 + 2 * 8}
REL_L2_GATE = 1e-4

WORKLOAD = dict(nz=3999, Nt=4096, N=1025, R_factor=15.0, begin=0.0, end=3999 * 5e-6,
                t_window=1.6e-12, energy_uJ=5.0, wavelength=800e-9, n2=3.2e-20, Ui_eV=9.0,
                rho_max=2.1e22, tau_c=1.27e-15, tau_r=150e-15, delta_t=85e-15, w0=100e-6,
                enable_ste=False, rho_t_stride=0, write_output=False, verbose=False)
ENVELOPE = dict(lens_f=0.05, noise=0.01, seed=2)
# the CPU-sized case (128 x 512, 800 fs window, 2 uJ collimated, 200 steps of 10 um): the
# mixed path's errors per output are reported at this shape too
SMALL = dict(WORKLOAD, nz=199, Nt=512, N=129, R_factor=10.0, end=1.99e-3, t_window=800e-15,
             energy_uJ=2.0)
SMALL_ENVELOPE = dict(lens_f=None, noise=0.01, seed=1)


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _integrator(precision, save_stride=10 ** 9):
    import rapport_amd as R
    kw = dict(WORKLOAD, save_stride=save_stride, precision=precision)
    return R.Integrator(R.Config(**kw), envelope=R.envelope_lens_noise(**ENVELOPE))


def run_gpu(it, steps, warmup):
    """Times `steps` pipeline steps of the integrator `it` three times: first as plain
    launches (the throughput), then with HIP event pairs bracketing every Hankel GEMM launch
    (its duration for the roofline; the event records themselves slow the step by about
    5 %), then with one record of every history after each step (the reference's default
    save_stride = 1)."""
    import torch
    import rapport_amd._lib as L
    for _ in range(warmup):
        it._pipeline_step(it.dz)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    gemm_ev = []
    orig = L.call

    def timed_call(name, *args):
        if name.startswith("fp_hankel_gemm"):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            orig(name, *args)
            b.record(stream)
            gemm_ev.append((a, b))
        else:
            orig(name, *args)

    import rapport_amd.integrator as integ
    import rapport_amd.mixed as mixed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        it._pipeline_step(it.dz)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    integ.call, mixed.call = timed_call, timed_call
    try:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            it._pipeline_step(it.dz)
        torch.cuda.synchronize()
        elapsed_ev = time.perf_counter() - t1
    finally:
        integ.call, mixed.call = orig, orig
    it.k_save = 0
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for i in range(steps):
        it._pipeline_step(it.dz)
        if it.k_save == it.n_saves:
            it.k_save = 0
        it._record(i * it.dz)
    torch.cuda.synchronize()
    elapsed_rec = time.perf_counter() - t2
    gemm_ms = sum(a.elapsed_time(b) for a, b in gemm_ev) / max(len(gemm_ev), 1)
    return dict(elapsed=elapsed, steps_per_s=steps / elapsed, ms_per_step=1e3 * elapsed / steps,
                ms_per_step_gemm_events=1e3 * elapsed_ev / steps,
                ms_per_step_recording=1e3 * elapsed_rec / steps,
                gemm_ms=gemm_ms, n_gemm=len(gemm_ev), capped_rows=int(it.t.status.item()),
                Nr=it.Nr, Nt=it.Nt, zero_cols=list(it._zc))


def _perturbed(it, eps, seed=99):
    """Multiply the loaded input field by (1 + eps xi), xi complex standard normal."""
    import numpy as np
    import torch
    u = it.u
    rng = np.random.default_rng(seed)
    z = torch.from_numpy(rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape)).to(u.device)
    it.u = u * (1 + eps * z)


def accuracy(acc_steps, full_steps=0, cond_eps=1e-8):
    """Mixed vs FP64 runs of the same input, fluence recorded every 40 steps.
    Gate: fluence / field / electron-density rel-L2 after acc_steps steps. With
    full_steps > acc_steps both runs continue to full_steps and the per-z fluence
    rel-L2 is reported, next to an FP64 run whose input is perturbed by
    (1 + cond_eps xi) -- the conditioning of the problem itself (cond_eps = 1e-8 is
    about the rounding of storing the input in complex64)."""
    import numpy as np
    def up(x):
        x = np.asarray(x)
        return x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)

    def rl2(a, b):
        return float(np.linalg.norm(up(a) - up(b)) / np.linalg.norm(up(b)))
    runs, fields = {}, {}
    for prec in ("fp64", "mixed"):
        it = _integrator(prec, save_stride=40)
        res = it.propagate(acc_steps)
        fields[prec] = it.u.cpu().numpy()
        if full_steps > acc_steps:
            res = it.propagate(full_steps - acc_steps)
        runs[prec] = res
        del it
    k = (acc_steps - 1) // 40 + 1                   # records inside the gate window
    rp, rm = runs["fp64"], runs["mixed"]
    out = dict(steps=acc_steps, fluence_rz=rl2(rm["fluence_rz"][:k], rp["fluence_rz"][:k]),
               fluence_final=rl2(rm["fluence_rz"][k - 1], rp["fluence_rz"][k - 1]),
               field=rl2(fields["mixed"], fields["fp64"]),
               rho_rz=rl2(rm["rho_rz"][:k], rp["rho_rz"][:k]),
               Imax_Wcm2=float(rp["Imax_z"][k - 1]))
    if full_steps > acc_steps:
        it = _integrator("fp64", save_stride=40)
        _perturbed(it, cond_eps)
        it.propagate(acc_steps)
        rc = it.propagate(full_steps - acc_steps)
        del it
        per_m = [rl2(rm["fluence_rz"][j], rp["fluence_rz"][j]) for j in range(rp["fluence_rz"].shape[0])]
        per_c = [rl2(rc["fluence_rz"][j], rp["fluence_rz"][j]) for j in range(rp["fluence_rz"].shape[0])]
        within = 0
        for j, e in enumerate(per_m):
            if e > REL_L2_GATE:
                break
            within = 40 * j
        out["full_run"] = dict(
            steps=full_steps, fluence_rz=rl2(rm["fluence_rz"], rp["fluence_rz"]),
            rho_rz=rl2(rm["rho_rz"], rp["rho_rz"]), last_step_within_gate=within,
            per_step_fluence={str(40 * j): per_m[j] for j in range(0, len(per_m), 5)},
            fp64_input_perturbed=dict(eps=cond_eps, fluence_rz=rl2(rc["fluence_rz"], rp["fluence_rz"]),
                                      per_step_fluence={str(40 * j): per_c[j]
                                                        for j in range(0, len(per_c), 5)}))
    return out


def _hybrid_step(it, stages):
    """One propagate() iteration of a mixed-precision Integrator with the named stages run
    in FP64 instead (on the complex64 state upcast to complex128, the result rounded back
    to complex64): "plasma" (inverse transform + carrier scan), "gemm" (the four Hankel
    GEMMs, FP64 MFMA with the FP64 half-step phase), "nl" (the nonlinear half of the Strang
    step). Stages not named run the mixed kernels. Used to find which stage's rounding
    drives the difference between the mixed and the FP64 runs."""
    import ctypes
    import torch
    from rapport_amd._lib import call, ptr, stream_ptr
    mx, m, t = it._mx, it._mx.m, it.t
    Nr, Nt, Np, dz = it.Nr, it.Nt, mx.Np, it.dz
    st = stream_ptr()
    c128 = torch.complex128
    if "plasma" in stages:
        s = mx.spec_rows.to(c128)
        call("fp_plasma_scan_spec", ptr(s), ptr(it.rho), None, None, ctypes.byref(t.plasma_params),
             ptr(t.W_lut), ptr(t.Ws_lut), ptr(t.tw), ptr(t.status), st)
    else:
        mx.plasma(it.rho, it.rho_s)
    ph32, ph64 = m.phase_half(dz), t.phase_half(dz)

    def gemm(B, C, alpha32, alpha64, rs32, rs64, with_phase):
        if "gemm" in stages:
            b = B[:Nr].to(c128)
            c = torch.empty_like(b)
            call("fp_hankel_gemm", ptr(t.Yt), ptr(b), ptr(c), Nr, Nt, alpha64, ptr(rs64),
                 ptr(ph64) if with_phase else None, st)
            C[:Nr].copy_(c)
        else:
            mx.slots.zero_()
            call("fp_absmax_f32", ptr(B), 2 * Np * Nt, mx._slot(0), st)
            call("fp_hankel_gemm_mixed", ptr(m.Yh), ptr(m.Yl), m.y_scale_exp, ptr(B), ptr(C), Np,
                 Nt, ctypes.c_float(alpha32), ptr(rs32), ptr(ph32) if with_phase else None,
                 mx._slot(0), None, None, st)
    gemm(mx.spec, mx.w1, m.Rk, t.Rk, None, None, True)
    gemm(mx.w1, mx.w0, m.iRk, t.iRk, None, None, False)
    if "nl" in stages:
        w = mx.w0[:Nr].to(c128)
        call("fp_nl_step", ptr(w), ptr(w), ptr(it.rho), None, 1, 1,
             ctypes.byref(it.nl.params(dz)), ctypes.byref(t.tables), st)
        mx.w0[:Nr].copy_(w)
    else:
        call("fp_nl_step_c64", ptr(mx.w0), ptr(mx.w0), ptr(it.rho), None, 1, 1,
             ctypes.byref(it.nl.params(dz)), ctypes.byref(m.tables), None, st)
    gemm(mx.w0, mx.w1, m.Rk, t.Rk, None, None, True)
    gemm(mx.w1, mx.spec, m.iRk, t.iRk, m.mask_r, t.mask_r, False)


def accuracy_sources(steps, every=40):
    """Where the mixed path's difference from FP64 comes from, over `steps` steps of the
    1024 x 4096 workload (fluence recorded every `every` steps, rel-L2 per record against
    the FP64 run):
    * floor: the FP64 pipeline with its state rounded to complex64 after every step -- the
      error every method that stores the field in complex64 starts from;
    * mixed with one stage at a time in FP64 (plasma, gemm, nl; see _hybrid_step), and
      with all three (then only the complex64 storage between the stages is left)."""
    import numpy as np
    import torch

    def up(x):
        return np.asarray(x).astype(np.float64)

    def curve(res, ref):
        return [float(np.linalg.norm(up(res[j]) - up(ref[j])) / np.linalg.norm(up(ref[j])))
                for j in range(ref.shape[0])]
    it = _integrator("fp64", save_stride=every)
    ref = it.propagate(steps)["fluence_rz"]
    del it
    out = {}
    it = _integrator("fp64", save_stride=every)
    for i in range(steps):
        it._pipeline_step(it.dz)
        it._spec.copy_(it._spec.to(torch.complex64).to(torch.complex128))
        if i % every == 0:
            it._record(i * it.dz)
    out["floor_fp64_c64_state"] = curve(it._results()["fluence_rz"], ref)
    del it
    variants = {"mixed": (), "plasma_fp64": ("plasma",), "gemm_fp64": ("gemm",),
                "nl_fp64": ("nl",), "all_stages_fp64": ("plasma", "gemm", "nl")}
    for name, stages in variants.items():
        it = _integrator("mixed", save_stride=every)
        for i in range(steps):
            _hybrid_step(it, stages)
            if i % every == 0:
                it._record(i * it.dz)
        out[name] = curve(it._results()["fluence_rz"], ref)
        del it
    torch.cuda.synchronize()
    pick = range(0, len(ref), max(1, 200 // every))
    return dict(steps=steps, every=every,
                final={k: v[-1] for k, v in out.items()},
                rms_over_run={k: float(np.sqrt(np.mean(np.square(v)))) for k, v in out.items()},
                per_step_fluence={k: {str(every * j): v[j] for j in pick} for k, v in out.items()})


def run_batched(n_inst=64, steps=20, seed=0, tune=True):
    """Batched sweep of n_inst instances (Nr in {256, 512, 768, 1024}, Nt in {1024, 2048,
    4096}, 0.5-20 uJ log-uniform, 40-200 fs, w0 50-150 um, 1 % noise, dz 10 um), tuned,
    one graph replay per z-step; per-pass times from HIP events around single passes."""
    import torch
    from rapport_amd.autotune import _time
    from rapport_amd.batch import BatchedIntegrator, random_sweep
    from rapport_amd._lib import call, stream_ptr
    cfgs, envs = random_sweep(n_inst, seed=seed)
    b = BatchedIntegrator(cfgs, envs, tune=tune)
    st = stream_ptr()
    passes = {}
    for name, p in (("plasma", 0), ("G1", 1), ("G2", 2), ("nl", 3), ("G3", 4), ("G4", 5)):
        passes[name] = _time(lambda: call("fp_batch_pass", b._plan, p, st), 1, 3)
    b.refresh_scales()
    g = b.capture(1)
    for _ in range(2):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    pts = b.points()
    gemm_ms = passes["G1"] + passes["G2"] + passes["G3"] + passes["G4"]
    ng_ms = passes["plasma"] + passes["nl"]
    ng_bytes = NONGEMM_BYTES_PER_PT["mixed"] * pts
    finite = all(bool(torch.isfinite(m._mx.spec).all()) for m in b.members)
    out = dict(instances=n_inst, steps=steps, ms_per_step=1e3 * dt, inst_steps_per_s=n_inst / dt,
               points=pts, gpts_steps_per_s=pts / dt / 1e9, launches_per_step=b.launches_per_step(),
               pass_ms=passes,
               roofline_nongemm=dict(bound="hbm", bytes_per_step=ng_bytes, ms_per_step=ng_ms,
                                     achieved=ng_bytes / (ng_ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS,
                                     unit="GB/s", frac=ng_bytes / (ng_ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
               gemm=dict(flops_per_step=b.gemm_flops(), ms_per_step=gemm_ms,
                         achieved_tflops=b.gemm_flops() / (gemm_ms * 1e-3) / 1e12,
                         peak_tflops=PEAK_TFLOPS["mixed"]),
               tuned={k: v for k, v in b.config.items() if k != "timings_ms"},
               finite=finite, capped_rows=sum(int(m.t.status.item()) for m in b.members))
    b.close()
    # the same sweep with one launch per pass (six per z-step)
    bm = BatchedIntegrator(cfgs, envs)
    bm.set_merged(True)
    passes_m = {}
    for name, p in (("plasma", 0), ("G1", 1), ("G2", 2), ("nl", 3), ("G3", 4), ("G4", 5)):
        passes_m[name] = _time(lambda: call("fp_batch_pass", bm._plan, p, st), 1, 3)
    bm.refresh_scales()
    gm = bm.capture(1)
    for _ in range(2):
        gm.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        gm.replay()
    torch.cuda.synchronize()
    dtm = (time.perf_counter() - t0) / steps
    out["single_launch_per_pass"] = dict(ms_per_step=1e3 * dtm, inst_steps_per_s=n_inst / dtm,
                                         launches_per_step=bm.launches_per_step(),
                                         pass_ms=passes_m)
    bm.close()
    del b, g, bm, gm
    torch.cuda.empty_cache()
    return out


ADJ_WORKLOAD = dict(nz=999, Nt=2048, N=513, R_factor=15.0, begin=0.0, end=999 * 10e-6,
                    energy_uJ=3.0, wavelength=800e-9, n2=3.2e-20, Ui_eV=9.0, rho_max=2.1e22,
                    tau_c=1.27e-15, tau_r=150e-15, delta_t=85e-15, w0=100e-6, enable_ste=False,
                    save_stride=10 ** 6, rho_t_stride=0, write_output=False, verbose=False)


def run_adjoint(n_fd=4, h=1e-4, P=16, seed=5, fit_iters=20):
    """Gradient of the fluence-mismatch loss w.r.t. 2P Laguerre-Gauss coefficients after
    1000 steps at 512 x 2048 (mixed forward + hand-written adjoint, sqrt(n)
    checkpoints), timed; checked on n_fd coefficients against central differences of
    the FP64 forward (h = 1e-4)."""
    import numpy as np
    import torch
    import rapport_amd as R
    from rapport_amd import adjoint as AD
    n = ADJ_WORKLOAD["nz"] + 1
    itm = R.Integrator(R.Config(**ADJ_WORKLOAD, precision="mixed"))
    it64 = R.Integrator(R.Config(**ADJ_WORKLOAD, precision="fp64"))
    eng = AD.AdjointEngine(itm)
    dev = itm.device
    G = AD.lg_modes(itm.g["rlist"], itm.cfg.w0, P).to(dev)
    tenv = AD.time_envelope(itm.g, itm.cfg, dev)
    w = AD.radial_weights(itm.g, dev)
    E0 = float(itm.cfg.E0)
    theta0, hidden = AD.lg_problem(P, seed=seed)
    field = lambda th: AD.lg_field(torch.as_tensor(th, device=dev), G, E0, tenv)  # noqa: E731
    with torch.no_grad():
        F_t = AD.forward_fluence(it64, field(hidden).to(torch.complex128), n).clone()
    theta = theta0 + 0.5 * (hidden - theta0)
    th = torch.tensor(theta, dtype=torch.float64, device=dev, requires_grad=True)
    AD.fluence_from_input(field(th), eng, n).sum().backward()      # warm-up
    th.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = AD.fluence_loss(AD.fluence_from_input(field(th), eng, n), F_t, w)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = th.grad.cpu().numpy()

    def L64(x):
        with torch.no_grad():
            return float(AD.fluence_loss(AD.forward_fluence(it64, field(x).to(torch.complex128), n), F_t, w))
    idx = list(range(n_fd // 2)) + list(range(P, P + n_fd - n_fd // 2))
    fd = []
    for k in idx:
        e = np.zeros(2 * P)
        e[k] = h
        fd.append((L64(theta + e) - L64(theta - e)) / (2 * h))
    fd = np.array(fd)
    ga = g[idx]
    loss64 = L64(theta)        # the FP64 forward's loss at the same input (forward precision)
    fit = None
    if fit_iters > 0:
        # the use: L-BFGS on the 2P coefficients from theta0 towards the FP64 target fluence
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        th_fit, hist = AD.fit_lg(eng, F_t, torch.tensor(theta0), n, iters=fit_iters, P=P)
        torch.cuda.synchronize()
        th_fit = th_fit.cpu().numpy()

        def dist(sl):
            return float(np.linalg.norm(th_fit[sl] - hidden[sl]) /
                         np.linalg.norm(theta0[sl] - hidden[sl]))
        # amplitude coefficients a_p and phase coefficients b_p apart: the output fluence
        # fixes the amplitude profile well but the input phase only weakly (b_0, a constant
        # phase, not at all), so the distance over all 2P mixes a recovery with a
        # near-null direction
        fit = dict(iters=fit_iters, evaluations=len(hist), seconds=time.perf_counter() - t3,
                   loss_start=hist[0], loss_end=hist[-1], loss_min=min(hist),
                   loss_end_fp64_forward=L64(th_fit),
                   coef_dist_to_hidden_rel=dist(slice(0, 2 * P)),
                   amp_coef_dist_to_hidden_rel=dist(slice(0, P)),
                   phase_coef_dist_to_hidden_rel=dist(slice(P + 1, 2 * P)),
                   note="distances relative to the start point's; phase excludes b_0")
    return dict(workload="512 x 2048, 1000 steps dz 10 um, 3 uJ 85 fs w0 100 um, 2P = %d LG coefficients" % (2 * P),
                forward_s=t1 - t0, backward_s=t2 - t1, grad_per_s=1.0 / (t2 - t0),
                zsteps_per_s_forward=n / (t1 - t0), zsteps_per_s_backward=n / (t2 - t1),
                loss=float(loss.detach()), loss_fp64_forward=loss64, fit=fit, fd_check=dict(h=h, coefficients=idx, adjoint=ga.tolist(),
                                                           fd_fp64=fd.tolist(),
                                                           rel_l2=float(np.linalg.norm(ga - fd) / np.linalg.norm(fd))))


def run_cpu_baseline(n_steps=5, repeats=3):
    """Oracle (NumPy FP64 restatement of the reference) on the host: one warm-up step,
    then `repeats` runs of n_steps timed steps of the same workload; the median run."""
    import statistics
    from oracle import reference_numpy as O
    kw = dict(WORKLOAD, save_stride=10 ** 9)
    it = O.Integrator(O.Config(**kw), envelope=O.envelope_lens_noise(**ENVELOPE))
    it.advance(1)
    rates, i = [], 2
    for _ in range(repeats):
        t0 = time.perf_counter()
        for _ in range(n_steps):
            it.advance(i)
            i += 1
        rates.append(n_steps / (time.perf_counter() - t0))
    return dict(value=statistics.median(rates), unit="z-steps/s", cores=_cores(), kind="port",
                runs=rates,
                sample=f"median of {repeats} runs of {n_steps} z-steps (after 1 warm-up) of the "
                       f"1024x4096 workload, NumPy/SciPy FP64 oracle, BLAS threads = {_cores()}")


def run_small_throughput(repeats=3, cpu_steps=20, gpu=True, cpu=True):
    """The CPU-sized case (128 x 512, 200 steps of 10 um, recording off) as a whole run:
    z-steps/s of the FP64 and mixed GPU paths (the 200 steps as graph replays from the same
    input, after one warm-up run; median of `repeats` runs) and of the NumPy FP64 oracle on the
    host (median of `repeats` runs of `cpu_steps` steps after one warm-up step)."""
    import statistics
    out = dict(shape="128 x 512", steps=200, recording="off")
    if gpu:
        import torch
        import rapport_amd as R
        n = SMALL["nz"] + 1
        for prec in ("fp64", "mixed"):
            kw = dict(SMALL, save_stride=10 ** 9, precision=prec)
            it = R.Integrator(R.Config(**kw), envelope=R.envelope_lens_noise(**SMALL_ENVELOPE))
            g = it.capture_step(1)
            snap = it._snapshot()
            rates = []
            for r in range(repeats + 1):
                it._restore(snap)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(n):
                    g.replay()
                torch.cuda.synchronize()
                if r > 0:
                    rates.append(n / (time.perf_counter() - t0))
            out[prec] = dict(value=statistics.median(rates), unit="z-steps/s", runs=rates)
            del g, it
        out["mixed_speedup_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    if cpu:
        from oracle import reference_numpy as O
        kw = dict(SMALL, save_stride=10 ** 9)
        it = O.Integrator(O.Config(**kw), envelope=O.envelope_lens_noise(**SMALL_ENVELOPE))
        it.advance(1)
        rates, i = [], 2
        for _ in range(repeats):
            t0 = time.perf_counter()
            for _ in range(cpu_steps):
                it.advance(i)
                i += 1
            rates.append(cpu_steps / (time.perf_counter() - t0))
        out["cpu_baseline"] = dict(value=statistics.median(rates), unit="z-steps/s",
                                   cores=_cores(), kind="port", runs=rates,
                                   sample=f"median of {repeats} runs of {cpu_steps} z-steps "
                                          "(after 1 warm-up) of the 128x512 run, NumPy/SciPy "
                                          f"FP64 oracle, BLAS threads = {_cores()}")
    return out


def copy_calibration(nbytes=256 * 2 ** 20, reps=5):
    """A plain device-to-device copy of a known size: the FETCH_SIZE / WRITE_SIZE
    counters of its kernel calibrate the counters read for the propagator kernels."""
    import torch
    x = torch.rand(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_bytes=nbytes, reps=reps)), flush=True)


def _pmc_traffic(kernel_substr, section="zstep_1024x4096"):
    """HBM bytes per launch of a kernel from the committed PMC summary (rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE runs, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        data = json.load(open(path))[section]
    except (OSError, KeyError, ValueError):
        return None
    for name, v in data.items():
        if kernel_substr in name:
            return dict(bytes=v["read_bytes"] + v["write_bytes"], read_bytes=v["read_bytes"],
                        write_bytes=v["write_bytes"], kernel=name,
                        source="profiles/pmc_traffic.json")
    return None


def _roofline(prec, r):
    Nr, Nt = r["Nr"], r["Nt"]
    # real Y (Nr x Nr) @ the 2 Nt real columns of the complex field, less the band of
    # columns where the spectral mask is exactly 0 (the band GEMMs skip it: those outputs
    # are 0 and are written, not computed)
    zlo, zhi = r.get("zero_cols", (0, 0))
    cols = 2 * Nt - (zhi - zlo)
    flops = 2.0 * Nr * Nr * cols
    ach = flops / (r["gemm_ms"] * 1e-3) / 1e12
    peak = PEAK_TFLOPS[prec]
    # the mixed step's G1 / G3 (phase epilogue) run the band tile, G2 / G4 the K-split tile:
    # gemm_ms is the average over the four launches, and so is the traffic
    knames = (["hankel_gemm_f64_band"] if prec == "fp64" else
              ["hankel_gemm_f16x3_band", "hankel_gemm_f16x3_ksplit"])
    trs = [_pmc_traffic(k) for k in knames]
    tr = None
    if all(trs):
        tr = {key: sum(t[key] for t in trs) / len(trs) for key in ("bytes", "read_bytes", "write_bytes")}
        tr.update(kernels=[t["kernel"] for t in trs], source=trs[0]["source"])
    kname = " + ".join(knames)
    roof = dict(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
                traffic=tr["bytes"] if tr else None, kernel=kname,
                flops_per_launch=flops, computed_real_columns=cols,
                dense_equivalent_flops_per_launch=4.0 * Nr * Nr * Nt,
                avg_launch_ms=r["gemm_ms"], launches_timed=r["n_gemm"])
    if tr:
        roof["traffic_detail"] = tr
    if prec == "mixed":
        roof["executed_flops_per_launch"] = 3 * flops
        roof["executed_frac"] = 3 * ach / peak
    rest_ms = r["ms_per_step_gemm_events"] - 4 * r["gemm_ms"]    # same run as gemm_ms
    nb = NONGEMM_BYTES_PER_PT[prec] * Nr * Nt
    ng = dict(bound="hbm", bytes_per_step=nb, ms_per_step=rest_ms,
              achieved=nb / (rest_ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s")
    ng["frac"] = ng["achieved"] / HBM_PEAK_GBS
    return roof, ng


def accuracy_small():
    """Mixed vs FP64 at 128 x 512 over the whole 200-step run (fluence recorded every
    step): rel-L2 of the final field, the electron density and the fluence."""
    import numpy as np
    import rapport_amd as R
    res, fields = {}, {}
    for prec in ("fp64", "mixed"):
        kw = dict(SMALL, save_stride=1, precision=prec)
        it = R.Integrator(R.Config(**kw), envelope=R.envelope_lens_noise(**SMALL_ENVELOPE))
        res[prec] = it.propagate()
        fields[prec] = it.u.cpu().numpy().astype(np.complex128)
        del it

    def rl2(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    rp, rm = res["fp64"], res["mixed"]
    return dict(shape="128 x 512", steps=200,
                field=float(np.linalg.norm(fields["mixed"] - fields["fp64"]) /
                            np.linalg.norm(fields["fp64"])),
                rho_rz=rl2(rm["rho_rz"], rp["rho_rz"]),
                fluence_rz=rl2(rm["fluence_rz"], rp["fluence_rz"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--acc-steps", type=int, default=400)
    ap.add_argument("--full-steps", type=int, default=4000,
                    help="continue the accuracy runs to this many steps (0: gate window only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sources-steps", type=int, default=4000,
                    help="steps of the error-source runs (floor and stage swaps; 0: skip)")
    ap.add_argument("--no-small-accuracy", action="store_true",
                    help="skip the 128 x 512 case (mixed-vs-FP64 check and whole-run z-steps/s)")
    ap.add_argument("--batched", type=int, default=64,
                    help="instances of the batched sweep (0: skip the batched measurement)")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    ap.add_argument("--fit-iters", type=int, default=20,
                    help="L-BFGS iterations of the input-profile fit (0: no fit)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the propagator is a single-GPU build); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if os.environ.get("RANK", "0") != "0":
        return

    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    # host set-up of every integrator first, so the timed runs follow one another without
    # the GPU idling in between
    its = {p: _integrator(p, save_stride=1) for p in precs}
    runs = {}
    for p in precs:
        runs[p] = run_gpu(its.pop(p), args.steps, args.warmup)
    acc = accuracy(args.acc_steps, args.full_steps) if args.precision == "both" else None
    head = "mixed" if (acc is not None and acc["fluence_rz"] <= REL_L2_GATE) or precs == ["mixed"] \
        else "fp64"
    r = runs[head]
    roof, ng = _roofline(head, r)
    out = dict(
        metric="z-steps/s at Nr x Nt = 1024 x 4096 (mixed vs FP64) at fluence rel-L2 <= 1e-4 "
               f"over {args.acc_steps} steps",
        value=r["steps_per_s"], unit="z-steps/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=r["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype=("c64 field / fp16x3 MFMA fp32-accum / f64 plasma" if head == "mixed" else "f64"),
        data="synthetic",
        config=dict(workload="fused-silica split-step propagator, 1024 radial rows x 4096 t-samples, "
                             "dz 5 um, 5 uJ 85 fs w0 100 um, lens f 5 cm, 1% seeded noise, STE off",
                    Nr=r["Nr"], Nt=r["Nt"], precision=head, recording="off"),
        roofline=roof, roofline_nongemm=ng,
        fp64=dict(value=runs["fp64"]["steps_per_s"], ms_per_step=runs["fp64"]["ms_per_step"],
                  ms_per_step_gemm_events=runs["fp64"]["ms_per_step_gemm_events"],
                  gemm_ms=runs["fp64"]["gemm_ms"],
                  ms_per_step_recording_every_step=runs["fp64"]["ms_per_step_recording"])
        if "fp64" in runs else None,
        mixed=dict(value=runs["mixed"]["steps_per_s"], ms_per_step=runs["mixed"]["ms_per_step"],
                   ms_per_step_gemm_events=runs["mixed"]["ms_per_step_gemm_events"],
                   gemm_ms=runs["mixed"]["gemm_ms"],
                   ms_per_step_recording_every_step=runs["mixed"]["ms_per_step_recording"])
        if "mixed" in runs else None,
        mixed_speedup_vs_fp64=(runs["mixed"]["steps_per_s"] / runs["fp64"]["steps_per_s"]
                               if len(runs) == 2 else None),
        accuracy_mixed_vs_fp64=acc,
        accuracy_mixed_vs_fp64_128x512=(accuracy_small() if args.precision == "both"
                                        and not args.no_small_accuracy else None),
        plasma_scan_capped_rows=sum(x["capped_rows"] for x in runs.values()),
    )
    if args.precision == "both" and args.sources_steps > 0:
        out["accuracy_sources"] = accuracy_sources(args.sources_steps)
    if args.batched > 0:
        out["batched"] = run_batched(args.batched, args.batched_steps)
    if not args.no_adjoint:
        out["adjoint"] = run_adjoint(fit_iters=args.fit_iters)
    if args.precision == "both" and not args.no_small_accuracy:
        out["small_128x512"] = run_small_throughput(cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = run_cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
