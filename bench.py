"""Benchmark: z-steps/s of the split-step propagator at Nr x Nt = 1024 x 4096.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed|both]
                    [--no-cpu-baseline]

Workload (synthetic input, host-generated and uploaded before timing): fused
silica at 800 nm, Nr = 1024 radial rows (Hankel order 1025, R = 1.5 mm),
Nt = 4096 samples over a 1.6 ps window, dz = 5 um, 5 uJ / 85 fs / w0 = 100 um
Gaussian with a f = 5 cm lens phase and 1 % seeded complex noise, STE off. A
step is one full iteration of the propagate() loop (plasma scan + Strang step)
with recording off; the field stays resident in HBM throughout.

Prints ONE JSON line. `roofline` is for the dominant kernel (the FP64 Hankel
GEMM, 4 per step), its average launch time measured with HIP events on the
launch stream inside the timed region; `cpu_baseline` times the NumPy FP64
oracle (a restatement of the reference algorithm) on the host for a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
FP16_MFMA_PEAK_TFLOPS = 2500.0

WORKLOAD = dict(nz=3999, Nt=4096, N=1025, R_factor=15.0, begin=0.0, end=3999 * 5e-6,
                t_window=1.6e-12, energy_uJ=5.0, wavelength=800e-9, n2=3.2e-20, Ui_eV=9.0,
                rho_max=2.1e22, tau_c=1.27e-15, tau_r=150e-15, delta_t=85e-15, w0=100e-6,
                enable_ste=False, rho_t_stride=0, write_output=False, verbose=False)
ENVELOPE = dict(lens_f=0.05, noise=0.01, seed=2)


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def run_gpu(precision, steps, warmup):
    import ctypes
    import torch
    import rapport_amd as R
    from rapport_amd._lib import call, ptr, stream_ptr

    kw = dict(WORKLOAD, save_stride=10 ** 9, precision=precision)
    it = R.Integrator(R.Config(**kw), envelope=R.envelope_lens_noise(**ENVELOPE))
    for _ in range(warmup):
        it._pipeline_step(it.dz)
    torch.cuda.synchronize()

    # time the steps, with event pairs around every Hankel GEMM launch on the launch stream
    t, Nr, Nt = it.t, it.Nr, it.Nt
    stream = torch.cuda.current_stream()
    gemm_ev = []
    orig_call = call

    def timed_call(name, *args):
        if name == "fp_hankel_gemm":
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            orig_call(name, *args)
            b.record(stream)
            gemm_ev.append((a, b))
        else:
            orig_call(name, *args)

    import rapport_amd.integrator as integ
    integ.call = timed_call
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            it._pipeline_step(it.dz)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    finally:
        integ.call = orig_call
    gemm_ms = sum(a.elapsed_time(b) for a, b in gemm_ev) / max(len(gemm_ev), 1)
    status = int(t.status.item())
    fl = (it.u.abs() ** 2).sum(dim=1).cpu().numpy()
    return dict(elapsed=elapsed, ms_per_step=1e3 * elapsed / steps, gemm_ms=gemm_ms,
                n_gemm=len(gemm_ev), status=status, fluence_rows=fl, Nr=Nr, Nt=Nt)


def run_cpu_baseline(n_steps=2):
    """Oracle (NumPy FP64 restatement of the reference) on the host: 1 warm-up
    step then n_steps timed steps of the same workload."""
    import numpy as np  # noqa: F401
    from oracle import reference_numpy as O
    kw = dict(WORKLOAD, save_stride=10 ** 9)
    it = O.Integrator(O.Config(**kw), envelope=O.envelope_lens_noise(**ENVELOPE))
    it.advance(1)
    t0 = time.perf_counter()
    for i in range(n_steps):
        it.advance(2 + i)
    dt = time.perf_counter() - t0
    return dict(value=n_steps / dt, unit="z-steps/s", cores=_cores(), kind="port",
                sample=f"{n_steps} z-steps (after 1 warm-up) of the 1024x4096 workload, NumPy/SciPy "
                       f"FP64 oracle, BLAS/FFT threads = {_cores()}")


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "both"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    r = run_gpu("fp64", args.steps, args.warmup)
    Nr, Nt = r["Nr"], r["Nt"]
    gemm_flops = 4.0 * Nr * Nr * Nt                     # real Y (Nr x Nr) @ complex (Nr x Nt)
    achieved = gemm_flops / (r["gemm_ms"] * 1e-3) / 1e12
    value = args.steps / r["elapsed"]
    out = dict(
        metric="z-steps/s at Nr x Nt = 1024 x 4096 (FP64 drop-in)",
        value=value, unit="z-steps/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=r["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic",
        config=dict(workload="fused-silica split-step propagator, 1024 radial rows x 4096 t-samples, "
                             "dz 5 um, 5 uJ 85 fs w0 100 um, lens f 5 cm, 1% seeded noise, STE off",
                    Nr=Nr, Nt=Nt, precision="fp64", recording="off"),
        roofline=dict(bound="mfma", achieved=achieved, peak=FP64_MFMA_PEAK_TFLOPS, unit="TFLOP/s",
                      frac=achieved / FP64_MFMA_PEAK_TFLOPS, traffic=None,
                      kernel="hankel_gemm_f64", flops_per_launch=gemm_flops,
                      avg_launch_ms=r["gemm_ms"], launches_timed=r["n_gemm"]),
        plasma_scan_capped_rows=r["status"],
    )
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = run_cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
