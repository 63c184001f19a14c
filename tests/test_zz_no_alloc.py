"""fp_record_history_noalloc (per-row partials in the library's static device array)
against fp_record_history (caller scratch): bitwise equal histories, on two calls in
a row, the second on a smaller field (fewer rows, shorter rows), for the FP64 and the
mixed-precision spectral state. Both entry points run the same kernels, so the bar is
exact equality of every output.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tests._cases import MATERIAL  # noqa: E402


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _integrator(N, Nt, precision):
    import rapport_amd as R
    kw = dict(MATERIAL, nz=3, Nt=Nt, N=N, R_factor=6.0, begin=0.0, end=3 * 3e-6,
              t_window=400e-15, energy_uJ=15.0, w0=30e-6, save_stride=1, rho_t_stride=3,
              rho_r_stride=2, absorb_time_bins_fs=[-100.0, 0.0, 100.0], rho_snapshot_t_fs=10.0,
              write_output=False, verbose=False, enable_ste=True)
    it = R.Integrator(R.Config(**kw, precision=precision),
                      envelope=R.envelope_lens_noise(noise=0.05, seed=3))
    for i in range(3):          # a state with plasma and STE populations
        it._pipeline_step(it.dz)
    torch.cuda.synchronize()
    return it


def _dst(it):
    """One record's worth of fresh destination buffers (NaN-filled) and its RecordDst."""
    from rapport_amd._lib import RecordDst
    dev = it.device
    Nr, Nt = it.Nr, it.Nt
    f32 = dict(dtype=torch.float32, device=dev)
    nb = int(it._bin_lo.numel())
    Nt_sub = (Nt - 1) // it._t_stride + 1
    Nr_sub = (Nr - 1) // it._r_stride + 1
    b = dict(fluence=torch.full((Nr,), np.nan, **f32), rho=torch.full((Nr,), np.nan, **f32),
             rho_s=torch.full((Nr,), np.nan, **f32), ipeak=torch.full((Nr,), np.nan, **f32),
             absorbed=torch.full((Nr,), np.nan, **f32), bins=torch.full((nb, Nr), np.nan, **f32),
             rho_at=torch.full((Nr,), np.nan, **f32), rho_onaxis=torch.full((Nt,), np.nan, **f32),
             rho_s_onaxis=torch.full((Nt,), np.nan, **f32),
             I_onaxis=torch.full((Nt,), np.nan, **f32),
             rho_cube=torch.full((Nr_sub, Nt_sub), np.nan, **f32),
             rho_s_cube=torch.full((Nr_sub, Nt_sub), np.nan, **f32),
             I_cube=torch.full((Nr_sub, Nt_sub), np.nan, **f32),
             sc=torch.full((3,), np.nan, dtype=torch.float64, device=dev))
    d = RecordDst(fluence=b["fluence"].data_ptr(), rho=b["rho"].data_ptr(),
                  rho_s=b["rho_s"].data_ptr(), ipeak=b["ipeak"].data_ptr(),
                  absorbed=b["absorbed"].data_ptr(), bins=b["bins"].data_ptr(),
                  rho_at=b["rho_at"].data_ptr(), it_snap=it._it_snap,
                  rho_onaxis=b["rho_onaxis"].data_ptr(),
                  rho_s_onaxis=b["rho_s_onaxis"].data_ptr(), I_onaxis=b["I_onaxis"].data_ptr(),
                  r_stride=it._r_stride, t_stride=it._t_stride, Nt_sub=Nt_sub,
                  rho_cube=b["rho_cube"].data_ptr(), rho_s_cube=b["rho_s_cube"].data_ptr(),
                  I_cube=b["I_cube"].data_ptr(), r_w=it._r_w_dev.data_ptr(),
                  partial=None, imax=b["sc"].data_ptr(), dedz_photo=b["sc"].data_ptr() + 8,
                  dedz_plasma=b["sc"].data_ptr() + 16)
    return b, d


def _record(it, name, partial=None):
    from rapport_amd._lib import call, ptr, stream_ptr
    b, d = _dst(it)
    if partial is not None:
        d.partial = partial.data_ptr()
    if it._mx is not None:
        spec, c64, tw = it._mx.spec, 1, it._mx.m.tw
    else:
        spec, c64, tw = it._spec, 0, it.t.tw
    call(name, ptr(spec), c64, ptr(tw), ptr(it.rho), ptr(it.rho_s),
         ctypes.byref(it.nl.params(it.dz)), ctypes.byref(it.t.tables), it.dt,
         int(it._bin_lo.numel()), ptr(it._bin_lo), ptr(it._bin_hi), ctypes.byref(d), stream_ptr())
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in b.items()}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_record_equals_record(precision):
    # second call smaller: 256 -> 128 rows and Nt 1024 -> 512, so the second record reads
    # only part of what the first left in the shared array
    for N, Nt in ((257, 1024), (129, 512)):
        _announce(f"no-alloc record vs record: {precision}, Nr {N - 1}, Nt {Nt}")
        it = _integrator(N, Nt, precision)
        partial = torch.full((3, it.Nr), np.nan, dtype=torch.float64, device=it.device)
        ref = _record(it, "fp_record_history", partial)
        got = _record(it, "fp_record_history_noalloc")
        for k in ref:
            assert np.all(np.isfinite(ref[k])), k
            assert np.array_equal(got[k], ref[k]), (k, N, Nt)
        assert float(np.max(ref["rho"])) > 0 and float(np.max(ref["rho_s"])) > 0
        del it
