"""Kernel-level parity of the HIP library against numpy / the CPU oracle.

Tolerances (float64 everywhere):
* FFT rows, Hankel GEMM: max relative error 1e-13 / 1e-12 of the row/matrix scale
  (different summation order than numpy's pocketfft / BLAS);
* plasma scan: rho within 1e-9 relative of the oracle's sequential recurrence
  (the reference's phi1 = (exp(x)-1)/x amplifies 1-ulp differences of exp by
  1/x near |x| = 1e-6, which caps attainable agreement well above 1e-16);
* split(), half_linear, fused nonlinear step: rel-L2 <= 1e-11.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import reference_numpy as O  # noqa: E402
from tests._cases import C1, C2, cfg_kwargs, envelope_args, pulse_field, rel_l2  # noqa: E402


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _setup(case, **over):
    import rapport_amd as R
    from rapport_amd.operators import DeviceTables
    kw = cfg_kwargs(case, **over)
    cfg_p = R.Config(**kw)
    cfg_o = O.Config(**kw)
    g_p = R.build_grids(cfg_p)
    g_o = O.build_grids(cfg_o)
    tabs = DeviceTables(cfg_p, g_p)
    return R, cfg_p, cfg_o, g_p, g_o, tabs


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("Nt", [512, 1024, 2048, 4096])
def test_fft_rows(Nt):
    from rapport_amd import _lib
    _announce(f"fft_rows Nt={Nt}")
    rng = np.random.default_rng(Nt)
    Nr = 64
    x = rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))
    tw = _cuda(np.exp(-2j * np.pi * np.arange(Nt) / Nt))
    xd = _cuda(x)
    yd = torch.empty_like(xd)
    _lib.call("fp_fft_rows", _lib.ptr(xd), _lib.ptr(yd), Nr, Nt, 0, _lib.ptr(tw), _lib.stream_ptr())
    ref = np.fft.fft(x, axis=1)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err
    _lib.call("fp_fft_rows", _lib.ptr(yd), _lib.ptr(xd), Nr, Nt, 1, _lib.ptr(tw), _lib.stream_ptr())
    err = np.abs(xd.cpu().numpy() - x).max() / np.abs(x).max()
    assert err < 1e-13, err


@pytest.mark.parametrize("Nt", [512, 1024, 2048, 4096])
def test_fft_rows_c64(Nt):
    """complex64 row transform (radix-8 passes, radix 16 at Nt = 4096) against NumPy's FP64
    transform of the same complex64 rows. The twiddles other than w come from products
    (w^2 = w w, w^4 = w^2 w^2, ...), within 15 x 2^-24 of exp(-2 pi i m / Nt); the bounds
    below are about 2.5x the measured FP32 rounding. A wrong or conjugated twiddle, a swapped
    output position or a missing 1/Nt gives errors of order 1."""
    from rapport_amd import _lib
    _announce(f"fft_rows_c64 Nt={Nt}")
    rng = np.random.default_rng(Nt + 1)
    Nr = 64
    x = (rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))).astype(np.complex64)
    tw = _cuda(np.exp(-2j * np.pi * np.arange(Nt) / Nt).astype(np.complex64))
    xd = _cuda(x)
    yd = torch.empty_like(xd)
    _lib.call("fp_fft_rows_c64", _lib.ptr(xd), _lib.ptr(yd), Nr, Nt, 0, _lib.ptr(tw),
              _lib.stream_ptr())
    ref = np.fft.fft(x.astype(np.complex128), axis=1)
    got = yd.cpu().numpy().astype(np.complex128)
    rms = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    mx = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"fft_rows_c64 Nt={Nt} forward rel-rms {rms:.2e} rel-max {mx:.2e}", flush=True)
    assert rms < 6e-7 and mx < 1.5e-6, (rms, mx)     # measured 1.7-3.0e-7 / 2.4-4.7e-7
    _lib.call("fp_fft_rows_c64", _lib.ptr(yd), _lib.ptr(xd), Nr, Nt, 1, _lib.ptr(tw),
              _lib.stream_ptr())
    back = xd.cpu().numpy().astype(np.complex128)
    x128 = x.astype(np.complex128)
    rms = np.linalg.norm(back - x128) / np.linalg.norm(x128)
    mx = np.abs(back - x128).max() / np.abs(x128).max()
    print(f"fft_rows_c64 Nt={Nt} round trip rel-rms {rms:.2e} rel-max {mx:.2e}", flush=True)
    assert rms < 1e-6 and mx < 1.5e-6, (rms, mx)     # measured 2.4-4.3e-7 / 3.1-5.9e-7


@pytest.mark.parametrize("Nr,Nt", [(128, 512), (200, 512), (1024, 4096)])
def test_hankel_gemm(Nr, Nt):
    from rapport_amd import _lib
    _announce(f"hankel_gemm Nr={Nr} Nt={Nt}")
    rng = np.random.default_rng(Nr)
    Y = rng.standard_normal((Nr, Nr))
    B = rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, (Nr, Nt)))
    rs = rng.uniform(0.5, 1.5, Nr)
    Yt, Bd, Cd = _cuda(Y.T), _cuda(B), torch.empty(Nr, Nt, dtype=torch.complex128, device="cuda")
    phd, rsd = _cuda(ph), _cuda(rs)
    ref = Y @ B
    _lib.call("fp_hankel_gemm", _lib.ptr(Yt), _lib.ptr(Bd), _lib.ptr(Cd), Nr, Nt, 0.75, None, None,
              _lib.stream_ptr())
    got = Cd.cpu().numpy()
    assert np.abs(got - 0.75 * ref).max() / np.abs(ref).max() < 1e-12
    _lib.call("fp_hankel_gemm", _lib.ptr(Yt), _lib.ptr(Bd), _lib.ptr(Cd), Nr, Nt, 2.0, _lib.ptr(rsd),
              _lib.ptr(phd), _lib.stream_ptr())
    got = Cd.cpu().numpy()
    ref2 = 2.0 * rs[:, None] * ref * ph
    assert np.abs(got - ref2).max() / np.abs(ref2).max() < 1e-12


@pytest.mark.parametrize("Nr,Nt,zc,expect_band", [(1024, 4096, (4096, 6048), True),
                                                  (512, 2048, (1000, 2900), True),
                                                  (768, 4096, (3500, 5000), True),
                                                  (200, 512, (300, 700), False),
                                                  (512, 2048, (0, 0), False),
                                                  (128, 512, (0, 0), False)])
def test_hankel_gemm_band_equals_full(Nr, Nt, zc, expect_band):
    # the banded FP64 GEMM (tiles only over the columns outside the masked range [zc)) gives
    # bitwise fp_hankel_gemm's values there, with and without phase / row scale; zero_fill
    # writes 0 in the masked columns, otherwise they keep their contents (16-column
    # alignment: the range's ends are computed). At 512 x 2048 the tiles are 2 column blocks
    # wide; with no range (0, 0) the whole width is tiled in tiles of 1 .. 4 column blocks
    # where 128 x 128 tiles would not fill the GPU. Nr = 200 is not tiled by 128 rows: the
    # entry falls back to the full GEMM.
    from rapport_amd import _lib
    _announce(f"hankel_gemm_band Nr={Nr} Nt={Nt} zero columns {zc}")
    rng = np.random.default_rng(Nr + 1)
    Y = rng.standard_normal((Nr, Nr))
    B = rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, (Nr, Nt)))
    rs = rng.uniform(0.5, 1.5, Nr)
    Yt, Bd = _cuda(Y.T), _cuda(B)
    phd, rsd = _cuda(ph), _cuda(rs)
    live = np.ones(2 * Nt, bool)
    live[zc[0]:zc[1]] = False
    for rsp, php in ((None, None), (rsd, phd)):
        full = torch.empty(Nr, Nt, dtype=torch.complex128, device="cuda")
        _lib.call("fp_hankel_gemm", _lib.ptr(Yt), _lib.ptr(Bd), _lib.ptr(full), Nr, Nt, 0.75,
                  _lib.ptr(rsp), _lib.ptr(php), _lib.stream_ptr())
        for zf in (0, 1):
            band = torch.full((Nr, Nt), complex(7.0, 7.0), dtype=torch.complex128, device="cuda")
            _lib.call("fp_hankel_gemm_band", _lib.ptr(Yt), _lib.ptr(Bd), _lib.ptr(band), Nr, Nt,
                      0.75, _lib.ptr(rsp), _lib.ptr(php), zc[0], zc[1], zf, _lib.stream_ptr())
            f = full.cpu().numpy().view(np.float64)
            b = band.cpu().numpy().view(np.float64)
            assert np.array_equal(b[:, live], f[:, live]), (zf, rsp is None)
            dead = b[:, ~live]
            fd = f[:, ~live]
            # a column of the range is computed (alignment, padding of the last tile -- the
            # pipeline's data are zero there -- or the full-GEMM fallback when the band would
            # not fill the GPU better), zeroed (zero_fill) or left untouched
            mark = 0.0 if zf else 7.0
            assert np.all((dead == mark) | (dead == fd))
            if expect_band:
                assert (dead == mark).any()


@pytest.mark.parametrize("case,ste,peaks", [
    (C1, False, (1e12, 3e13, 6e13, 9e13)),
    (C1, True, (1e12, 3e13, 6e13, 9e13)),
    (C1, False, (2e14, 1e15, 5e15, 1e16)),       # depletion and the n_a clamp
    (C2, False, (1e12, 2e13, 5e13, 8e13)),
])
def test_plasma_scan_vs_oracle(case, ste, peaks):
    _announce(f"plasma_scan Nt={case['Nt']} ste={ste} peaks={peaks[-1]:.0e}")
    R, cfg_p, cfg_o, g_p, g_o, tabs = _setup(case, enable_ste=ste, tau_ste=1e-12 if ste else None)
    Nr, Nt = g_p["Y"].shape[0], cfg_p.Nt
    pk = np.resize(np.asarray(peaks), Nr)
    u = pulse_field(Nr, Nt, g_p["dt"], g_p["invE2"], pk, seed=5)
    nl_o = O.NonlinearOperator(cfg_o, g_o)
    rho_o, rhos_o = np.zeros((Nr, Nt)), np.zeros((Nr, Nt))
    nl_o.update_plasma(u, rho_o, rhos_o, g_o["dt"])
    nl_p = R.NonlinearOperator(cfg_p, g_p, tabs)
    ud = _cuda(u)
    rho = torch.full((Nr, Nt), 7.0, dtype=torch.float64, device="cuda")
    rhos = torch.full((Nr, Nt), 7.0, dtype=torch.float64, device="cuda")
    tabs.status.zero_()
    nl_p.update_plasma(ud, rho, rhos, g_p["dt"])
    got, gots = rho.cpu().numpy(), rhos.cpu().numpy()
    assert int(tabs.status.item()) == 0
    scale = np.maximum(np.abs(rho_o), 1e-30 * rho_o.max())
    err = np.max(np.abs(got - rho_o) / np.maximum(scale, 1e-12 * rho_o.max()))
    assert rho_o.max() > 1e15
    assert err < 1e-9, err
    if ste:
        assert np.abs(gots - rhos_o).max() <= 1e-9 * rhos_o.max()
    else:
        assert np.all(gots == 0)


@pytest.mark.parametrize("ste", [False, True])
def test_split_vs_oracle(ste):
    _announce(f"split C1 ste={ste}")
    R, cfg_p, cfg_o, g_p, g_o, tabs = _setup(C1, enable_ste=ste)
    Nr, Nt = g_p["Y"].shape[0], cfg_p.Nt
    u = pulse_field(Nr, Nt, g_p["dt"], g_p["invE2"], np.linspace(1e12, 7e13, Nr), seed=9)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0, 3e20, (Nr, Nt))
    rhos = rng.uniform(0, 1e20, (Nr, Nt)) if ste else None
    nl_o = O.NonlinearOperator(cfg_o, g_o)
    rhs_o, a_o = nl_o.split(u, rho, rhos)
    nl_p = R.NonlinearOperator(cfg_p, g_p, tabs)
    rhs, a = nl_p.split(_cuda(u), _cuda(rho), _cuda(rhos) if ste else None)
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < 1e-11
    assert rel_l2(a.cpu().numpy(), a_o) < 1e-12
    ph, pl = nl_p.loss_rates(_cuda(u), _cuda(rho))
    ph_o, pl_o = nl_o.loss_rates(u, rho)
    assert rel_l2(ph.cpu().numpy(), ph_o) < 1e-12
    assert rel_l2(pl.cpu().numpy(), pl_o) < 1e-14


def test_split_raman_only_vs_oracle():
    """Raman term alone: the device's recursion scan against the oracle's
    ifft(fft(I) R_f) form of the reference."""
    _announce("split C1 raman only")
    only = dict(enable_kerr_instantaneous=False, enable_kerr_raman=True,
                enable_photoionization_loss=False, enable_plasma_defocusing=False,
                enable_plasma_absorption=False, enable_ste=False)
    R, cfg_p, cfg_o, g_p, g_o, tabs = _setup(C1, **only)
    Nr, Nt = g_p["Y"].shape[0], cfg_p.Nt
    u = pulse_field(Nr, Nt, g_p["dt"], g_p["invE2"], np.linspace(1e12, 7e13, Nr), seed=4)
    rho = np.zeros((Nr, Nt))
    rhs_o, _ = O.NonlinearOperator(cfg_o, g_o).split(u, rho)
    rhs, _ = R.NonlinearOperator(cfg_p, g_p, tabs).split(_cuda(u), _cuda(rho))
    assert np.abs(rhs_o).max() > 0
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < 1e-11


def test_split_toggles_off_is_zero():
    _announce("split toggles off")
    off = dict(enable_kerr_instantaneous=False, enable_kerr_raman=False, enable_self_steepening=False,
               enable_photoionization_loss=False, enable_plasma_defocusing=False,
               enable_plasma_absorption=False)
    R, cfg_p, cfg_o, g_p, g_o, tabs = _setup(C1, **off)
    Nr, Nt = g_p["Y"].shape[0], cfg_p.Nt
    u = pulse_field(Nr, Nt, g_p["dt"], g_p["invE2"], np.full(Nr, 5e13), seed=1)
    rhs, a = R.NonlinearOperator(cfg_p, g_p, tabs).split(_cuda(u), _cuda(np.full((Nr, Nt), 1e19)))
    assert torch.count_nonzero(rhs).item() == 0 and torch.count_nonzero(a).item() == 0


@pytest.mark.parametrize("case", [C1, C2])
def test_half_linear_vs_oracle(case):
    _announce(f"half_linear Nt={case['Nt']}")
    R, cfg_p, cfg_o, g_p, g_o, tabs = _setup(case)
    u0 = O.envelope_lens_noise(**envelope_args(case))(g_o["rr"], g_o["tt"], cfg_o, g_o)
    ref = O.LinearOperator(cfg_o, g_o).half_linear(u0, cfg_o.dz)
    got = R.LinearOperator(cfg_p, g_p, tabs).half_linear(_cuda(u0), cfg_p.dz).cpu().numpy()
    assert rel_l2(got, ref) < 1e-11


@pytest.mark.parametrize("spectral", [False, True])
def test_nl_step_vs_oracle(spectral):
    from rapport_amd import _lib
    _announce(f"nl_step C1 spectral={spectral}")
    R, cfg_p, cfg_o, g_p, g_o, tabs = _setup(C1, energy_uJ=60.0)
    Nr, Nt = g_p["Y"].shape[0], cfg_p.Nt
    u = O.envelope_lens_noise(**envelope_args(C1))(g_o["rr"], g_o["tt"], cfg_o, g_o)
    rho, rhos = np.zeros((Nr, Nt)), np.zeros((Nr, Nt))
    nl_o = O.NonlinearOperator(cfg_o, g_o)
    nl_o.update_plasma(u, rho, rhos, g_o["dt"])
    dz = cfg_o.dz
    v = u * np.exp(-0.5 * dz * nl_o.alpha_only(u, rho))
    k1, _ = nl_o.split(v, rho, rhos)
    k2, _ = nl_o.split(v + 0.5 * dz * k1, rho, rhos)
    k3, _ = nl_o.split(v + 0.5 * dz * k2, rho, rhos)
    k4, _ = nl_o.split(v + dz * k3, rho, rhos)
    v = v + dz / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    ref = v * np.exp(-0.5 * dz * nl_o.alpha_only(v, rho))
    assert rho.max() > 1e15
    nl_p = R.NonlinearOperator(cfg_p, g_p, tabs)
    src = np.fft.fft(u, axis=1) if spectral else u
    ud = _cuda(src)
    out = torch.empty_like(ud)
    _lib.call("fp_nl_step", _lib.ptr(ud), _lib.ptr(out), _lib.ptr(_cuda(rho)), None,
              int(spectral), int(spectral), ctypes.byref(nl_p.params(dz)),
              ctypes.byref(tabs.tables), _lib.stream_ptr())
    got = out.cpu().numpy()
    if spectral:
        got = np.fft.ifft(got, axis=1)
    assert rel_l2(got, ref) < 1e-11
    assert rel_l2(got, u) > 1e-6          # the nonlinear step did something
