"""Mixed-precision path (complex64 field, split-FP16 Hankel GEMM with FP32
accumulation, FP32 transforms and RK4, FP64 carrier scan) against the FP64
references. Tolerances, stated per output:
* split-FP16 GEMM: max error <= 4e-6 of max|Y B| (FP32 accumulation over K);
* split(): rel-L2 <= 3e-5 against the oracle;
* whole runs: fluence rel-L2 <= 1e-4 against the FP64 GPU path (itself pinned to
  the oracle at 1e-9); field and electron density rel-L2 reported and bounded.
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


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("case", [C1, C2])
def test_split_fp16_gemm(case):
    import rapport_amd as R
    from rapport_amd import _lib
    from rapport_amd.mixed import split_hankel_fp16
    _announce(f"gemm_mixed N={case['N']} Nt={case['Nt']}")
    cfg = R.Config(**cfg_kwargs(case))
    Y = R.build_grids(cfg)["Y"]
    Nr, Nt = Y.shape[0], cfg.Nt
    rng = np.random.default_rng(11)
    B = (rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))) * 3e9
    B[: Nr // 8] *= 1e4                      # dynamic range across rows
    Bd = _cuda(B.astype(np.complex64))
    Cd = torch.empty_like(Bd)
    yh, yl, se = split_hankel_fp16(Y)
    Yh, Yl = _cuda(yh), _cuda(yl)
    slots = torch.zeros(4, dtype=torch.int32, device="cuda")
    sp = lambda i: ctypes.c_void_p(slots.data_ptr() + 4 * i)  # noqa: E731
    _lib.call("fp_absmax_f32", _lib.ptr(Bd), 2 * Nr * Nt, sp(0), _lib.stream_ptr())
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, (Nr, Nt)))
    phd = _cuda(ph.astype(np.complex64))
    _lib.call("fp_hankel_gemm_mixed", _lib.ptr(Yh), _lib.ptr(Yl), se, _lib.ptr(Bd), _lib.ptr(Cd),
              Nr, Nt, ctypes.c_float(0.5), None, _lib.ptr(phd), sp(0), sp(1), sp(2), _lib.stream_ptr())
    got = Cd.cpu().numpy().astype(np.complex128)
    ref = 0.5 * (Y @ B.astype(np.complex64).astype(np.complex128)) * ph.astype(np.complex64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 4e-6, err
    # the epilogue left max|C| in slot 1 and cleared slot 2
    s = slots.cpu().numpy().view(np.float32)
    assert s[1] == pytest.approx(np.abs(got.view(np.float64)).max(), rel=1e-6)
    assert slots[2].item() == 0


def test_split_c64_vs_oracle():
    import rapport_amd as R
    from rapport_amd import _lib
    from rapport_amd.mixed import MixedTables
    from rapport_amd.operators import DeviceTables
    _announce("split c64 C1")
    kw = cfg_kwargs(C1)
    cfg_p, cfg_o = R.Config(**kw), O.Config(**kw)
    g_p, g_o = R.build_grids(cfg_p), O.build_grids(cfg_o)
    t64 = DeviceTables(cfg_p, g_p)
    mt = MixedTables(t64, cfg_p, g_p)
    Nr, Nt = g_p["Y"].shape[0], cfg_p.Nt
    u = pulse_field(Nr, Nt, g_p["dt"], g_p["invE2"], np.linspace(1e12, 7e13, Nr), seed=9)
    rho = np.random.default_rng(3).uniform(0, 3e20, (Nr, Nt))
    rhs_o, a_o = O.NonlinearOperator(cfg_o, g_o).split(u, rho, None)
    ud = _cuda(u.astype(np.complex64))
    rhs = torch.empty_like(ud)
    al = torch.empty(Nr, Nt, dtype=torch.float32, device="cuda")
    _lib.call("fp_nl_split_c64", _lib.ptr(ud), _lib.ptr(_cuda(rho)), None, _lib.ptr(rhs), _lib.ptr(al),
              ctypes.byref(t64.nl_params(cfg_p.dz)), ctypes.byref(mt.tables), _lib.stream_ptr())
    assert rel_l2(rhs.cpu().numpy(), rhs_o) < 3e-5
    assert rel_l2(al.cpu().numpy(), a_o) < 3e-5


@pytest.mark.parametrize("Nt", [2048, 4096])
def test_nl_step_c64_wide_rows_vs_oracle(Nt):
    """The FP32 fused nonlinear step (fp_nl_step_c64) at the wide row shapes: 4096 samples
    run on 256 threads of 16 samples (three radix-16 passes per transform), 2048 samples on
    256 threads of 8. (a) Against the oracle's FP64 step (alpha/2, RK4 of split(), alpha/2)
    from the same complex64 spectral input on 8 ionising rows (about 1 rad of nonlinear phase
    per step): the row within 2e-5 and the step's increment within 3e-5 (measured 1.2e-5 and
    9e-6 at 4096: FP32 rounding of 18 transforms, the PCHIP and the RK4 sums). (b) With every term
    off the step is the round trip of its two outer transforms: the coherent gain
    Re<in, out> / <in, in> - 1 stays below 1e-8 (measured 5e-10 and 1e-9: the exact
    transforms; the plain float transform, with rounded rotation constants and twiddle powers
    formed by products, shrinks the field by about 1e-7 per step)."""
    import rapport_amd as R
    from rapport_amd import _lib
    from rapport_amd.mixed import MixedTables
    from rapport_amd.operators import DeviceTables
    from scipy.special import erf
    _announce(f"nl_step c64 Nt={Nt} vs oracle, round-trip gain")
    kw = cfg_kwargs(C1, Nt=Nt, N=9)
    cfg_p, cfg_o = R.Config(**kw), O.Config(**kw)
    g_p, g_o = R.build_grids(cfg_p), O.build_grids(cfg_o)
    t64 = DeviceTables(cfg_p, g_p)
    mt = MixedTables(t64, cfg_p, g_p)
    Nr = g_p["Y"].shape[0]
    assert Nr == 8
    # chirped 60 fs Gaussians with 1 % complex noise, band-limited by the spectral mask as the
    # propagator's fields are (white noise up to the Nyquist frequency, or a random phase per
    # sample, would make the oracle's own step unstable at this dz)
    t = (np.arange(Nt) - Nt / 2) * g_p["dt"]
    rng = np.random.default_rng(5)
    peaks = np.linspace(2e12, 3e13, Nr)[:, None]
    noise = 1 + 0.01 * (rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt)))
    u = np.sqrt(peaks / g_p["invE2"]) * np.exp(-(t / 60e-15) ** 2 + 1j * (t / 40e-15) ** 2) * noise
    spec = (np.fft.fft(u, axis=1) * g_o["spec_mask"][None, :]).astype(np.complex64)
    u = np.fft.ifft(spec.astype(np.complex128), axis=1)   # the oracle starts from the same input
    # a smooth density ramp (a density drawn per sample would also be white noise)
    rho = np.linspace(1e19, 3e20, Nr)[:, None] * 0.5 * (1 + erf(t / 50e-15))[None, :]
    nl_o = O.NonlinearOperator(cfg_o, g_o)
    dz = cfg_o.dz
    v = u * np.exp(-0.5 * dz * nl_o.alpha_only(u, rho))
    k1, _ = nl_o.split(v, rho, None)
    k2, _ = nl_o.split(v + 0.5 * dz * k1, rho, None)
    k3, _ = nl_o.split(v + 0.5 * dz * k2, rho, None)
    k4, _ = nl_o.split(v + dz * k3, rho, None)
    v = v + dz / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    ref = v * np.exp(-0.5 * dz * nl_o.alpha_only(v, rho))

    sd = _cuda(spec)
    out = torch.empty_like(sd)
    p = t64.nl_params(dz)
    _lib.call("fp_nl_step_c64", _lib.ptr(sd), _lib.ptr(out), _lib.ptr(_cuda(rho)), None, 1, 1,
              ctypes.byref(p), ctypes.byref(mt.tables), None, _lib.stream_ptr())
    got = np.fft.ifft(out.cpu().numpy().astype(np.complex128), axis=1)
    e_row, e_inc = rel_l2(got, ref), rel_l2(got - u, ref - u)
    print(f"[nl c64 Nt={Nt}] rel-L2 row {e_row:.2e} increment {e_inc:.2e} "
          f"(increment / row {rel_l2(ref, u):.2e})", flush=True)
    assert rel_l2(ref, u) > 0.1           # the step changes the rows (about 1 rad of phase)
    assert e_row < 2e-5 and e_inc < 3e-5

    q = t64.nl_params(dz)
    q.en_kerr_inst = q.en_kerr_raman = q.en_pi_loss = q.en_plasma_absorb = 0
    q.pc_re = q.pc_im = q.ste_pref = 0.0
    rng = np.random.default_rng(6)
    x = (rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))).astype(np.complex64)
    xd = _cuda(x)
    y = torch.empty_like(xd)
    _lib.call("fp_nl_step_c64", _lib.ptr(xd), _lib.ptr(y), _lib.ptr(_cuda(rho)), None, 1, 1,
              ctypes.byref(q), ctypes.byref(mt.tables), None, _lib.stream_ptr())
    yo = y.cpu().numpy().astype(np.complex128)
    xo = x.astype(np.complex128)
    gain = float(np.real(np.vdot(xo, yo)) / np.real(np.vdot(xo, xo)) - 1.0)
    print(f"[nl c64 Nt={Nt}] round trip rel-L2 {rel_l2(yo, xo):.2e} coherent gain {gain:.2e}",
          flush=True)
    assert rel_l2(yo, xo) < 1e-6
    assert abs(gain) < 1e-8


def _runs(case, steps, save_stride, **over):
    import rapport_amd as R
    env = R.envelope_lens_noise(**envelope_args(case))
    out = {}
    for prec in ("fp64", "mixed"):
        kw = cfg_kwargs(case, save_stride=save_stride, precision=prec, **over)
        it = R.Integrator(R.Config(**kw), envelope=env)
        res = it.propagate(steps)
        out[prec] = (it, res)
    return out


def test_mixed_c1_full_run_vs_fp64():
    _announce("mixed vs fp64, C1 200 steps")
    out = _runs(C1, None, 10)
    (ip, rp), (im, rm) = out["fp64"], out["mixed"]
    e_f = rel_l2(rm["fluence_rz"], rp["fluence_rz"])
    e_u = rel_l2(im.u.cpu().numpy(), ip.u.cpu().numpy())
    e_r = rel_l2(rm["rho_rz"], rp["rho_rz"])
    print(f"[mixed C1] rel-L2 fluence {e_f:.2e} field {e_u:.2e} rho {e_r:.2e}")
    assert e_f < 1e-4 and e_u < 1e-3 and e_r < 1e-3


def test_mixed_hot_run_vs_fp64():
    _announce("mixed vs fp64, hot 60 steps")
    hot = dict(C1, nz=59, end=59 * 10e-6, energy_uJ=20.0)
    out = _runs(hot, None, 5, w0=30e-6, R_factor=20.0)
    (ip, rp), (im, rm) = out["fp64"], out["mixed"]
    assert rp["rho_rz"].max() > 1e17
    e_f = rel_l2(rm["fluence_rz"], rp["fluence_rz"])
    e_r = rel_l2(rm["rho_rz"], rp["rho_rz"])
    e_u = rel_l2(im.u.cpu().numpy(), ip.u.cpu().numpy())
    print(f"[mixed hot] rel-L2 fluence {e_f:.2e} field {e_u:.2e} rho {e_r:.2e}")
    assert e_f < 1e-4 and e_r < 1e-3 and e_u < 1e-3


def test_ragged_radial_grid_n201_vs_oracle():
    # Nr = 200 (N = 201) is not a multiple of the GEMM's 128-row tile: the mixed engine pads
    # its operands to 256 rows with zeros. Hot 40 steps: the FP64 path against the oracle at
    # 1e-9, the mixed path against both at its 1e-4 fluence bar; padded rows stay exactly 0.
    import rapport_amd as R
    _announce("ragged radial grid N=201: fp64 vs oracle, mixed vs fp64 / oracle, 40 steps")
    hot = dict(C1, N=201, nz=39, end=39 * 10e-6, energy_uJ=6.0)   # 8-12 uJ overflow (reference too)
    over = dict(w0=30e-6, R_factor=20.0, save_stride=4)
    env = envelope_args(hot)
    kw = cfg_kwargs(hot, **over)
    ito = O.Integrator(O.Config(**kw), envelope=O.envelope_lens_noise(**env))
    ro = ito.propagate()
    out = _runs(hot, None, 4, **{k: v for k, v in over.items() if k != "save_stride"})
    (ip, rp), (im, rm) = out["fp64"], out["mixed"]
    assert ip.Nr == 200 and im._mx.Np == 256
    assert ro["rho_rz"].max() > 1e17 and np.isfinite(ito.u).all()
    for key in ("fluence_rz", "rho_rz", "Imax_z"):
        assert rel_l2(rp[key], ro[key]) < 1e-9, key
    assert rel_l2(ip.u.cpu().numpy(), ito.u) < 1e-9
    e_f = rel_l2(rm["fluence_rz"], ro["fluence_rz"])
    e_r = rel_l2(rm["rho_rz"], ro["rho_rz"])
    e_u = rel_l2(im.u.cpu().numpy(), ito.u)
    _announce(f"  mixed vs oracle: fluence {e_f:.2e} field {e_u:.2e} rho {e_r:.2e}")
    assert e_f < 1e-4 and e_r < 1e-3 and e_u < 1e-3
    pad = im._mx.spec[im.Nr:]
    assert pad.numel() > 0 and not bool(pad.abs().max() > 0)


def test_reference_production_grid_n3001_mixed_vs_fp64():
    # the reference's production radial grid (N = 3001, i.e. 3000 rows, padded to 3072 in
    # the mixed engine; filament_visualization_1030nm.ipynb:169) at Nt = 1024, 6 hot steps
    import rapport_amd as R
    _announce("reference radial grid N=3001 (3000 rows) x Nt 1024: mixed vs fp64, 6 steps")
    case = dict(C1, N=3001, Nt=1024, nz=5, end=5 * 5e-6, energy_uJ=6.0)
    out = _runs(case, None, 1, w0=30e-6, R_factor=60.0)
    (ip, rp), (im, rm) = out["fp64"], out["mixed"]
    assert im.Nr == 3000 and im._mx.Np == 3072
    e_f = rel_l2(rm["fluence_rz"], rp["fluence_rz"])
    e_r = rel_l2(rm["rho_rz"], rp["rho_rz"])
    e_u = rel_l2(im.u.cpu().numpy(), ip.u.cpu().numpy())
    _announce(f"  fluence {e_f:.2e} field {e_u:.2e} rho {e_r:.2e}, max rho {rp['rho_rz'].max():.2e}")
    assert rp["rho_rz"].max() > 1e15 and bool(torch.isfinite(ip.u).all())
    assert e_f < 1e-4 and e_r < 1e-3 and e_u < 1e-3


@pytest.mark.parametrize("case", [C2, dict(C2, N=769), dict(C2, N=513)])
def test_band_gemm_equals_full_gemm_outside_the_mask(case):
    # the banded single-instance GEMM (tiles only over the frequencies where the spectral mask
    # is not exactly 0: 7-, 5- and 4-block tiles at Nr = 1024, 768, 512) gives bitwise the
    # full GEMM's values there; with zero_fill the masked columns are written as 0, without
    # it they are left as they were.
    import rapport_amd as R
    from rapport_amd import _lib
    from rapport_amd.mixed import split_hankel_fp16, zero_columns
    _announce(f"band GEMM vs full GEMM, N={case['N']} Nt={case['Nt']}")
    cfg = R.Config(**cfg_kwargs(case))
    g = R.build_grids(cfg)
    Y = g["Y"]
    Nr, Nt = Y.shape[0], cfg.Nt
    zlo, zhi = zero_columns(g["spec_mask"])
    assert zhi - zlo > 1000
    rng = np.random.default_rng(5)
    B = (rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))) * 3e9
    Bd = _cuda(B.astype(np.complex64))
    yh, yl, se = split_hankel_fp16(Y)
    Yh, Yl = _cuda(yh), _cuda(yl)
    ph = _cuda((np.exp(1j * rng.uniform(0, 2 * np.pi, (Nr, Nt))) * g["spec_mask"][None, :])
               .astype(np.complex64))
    slots = torch.zeros(4, dtype=torch.int32, device="cuda")
    sp = lambda i: ctypes.c_void_p(slots.data_ptr() + 4 * i)  # noqa: E731
    st = _lib.stream_ptr()
    _lib.call("fp_absmax_f32", _lib.ptr(Bd), 2 * Nr * Nt, sp(0), st)
    full = torch.full_like(Bd, complex(7.0, 7.0))
    _lib.call("fp_hankel_gemm_mixed", _lib.ptr(Yh), _lib.ptr(Yl), se, _lib.ptr(Bd), _lib.ptr(full),
              Nr, Nt, ctypes.c_float(0.5), None, _lib.ptr(ph), sp(0), sp(1), None, st)
    for zf in (0, 1):
        band = torch.full_like(Bd, complex(7.0, 7.0))
        _lib.call("fp_hankel_gemm_mixed_band", _lib.ptr(Yh), _lib.ptr(Yl), se, _lib.ptr(Bd),
                  _lib.ptr(band), Nr, Nt, ctypes.c_float(0.5), None, _lib.ptr(ph), sp(0), sp(2),
                  None, zlo, zhi, zf, st)
        torch.cuda.synchronize()
        f = torch.view_as_real(full).reshape(Nr, 2 * Nt)
        b = torch.view_as_real(band).reshape(Nr, 2 * Nt)
        live = torch.ones(2 * Nt, dtype=torch.bool, device="cuda")
        live[zlo:zhi] = False
        assert torch.equal(b[:, live], f[:, live]), zf
        dead = b[:, ~live]
        if zf:
            assert bool((dead == 0).all())
        else:
            # 16-column alignment: the ends of the masked range may be computed (zeros)
            assert bool(((dead == 7.0) | (dead == 0)).all()) and bool((dead == 7.0).any())
        assert slots[2].item() == slots[1].item()   # same max|C|


@pytest.mark.parametrize("case", [C2, dict(C2, N=769), dict(C2, N=513)])
def test_band_gemm_without_phase_k_split(case):
    # without a phase (G2 / G4) the banded GEMM runs the K-split tile (two interleaved
    # partial sums over K, added at the end): the same products as the full GEMM in another
    # summation order, so the two agree to FP32 accumulation error, not bitwise. Both are
    # checked against Y @ B in FP64 (4e-6 of max|Y B|, the split-GEMM bound above), with a
    # row scale, the masked columns zero-filled, and max|C| reduced into its slot.
    import rapport_amd as R
    from rapport_amd import _lib
    from rapport_amd.mixed import split_hankel_fp16, zero_columns
    _announce(f"K-split band GEMM vs full GEMM and FP64, N={case['N']} Nt={case['Nt']}")
    cfg = R.Config(**cfg_kwargs(case))
    g = R.build_grids(cfg)
    Y = g["Y"]
    Nr, Nt = Y.shape[0], cfg.Nt
    zlo, zhi = zero_columns(g["spec_mask"])
    rng = np.random.default_rng(6)
    B = (rng.standard_normal((Nr, Nt)) + 1j * rng.standard_normal((Nr, Nt))) * 3e9
    B[: Nr // 8] *= 1e4
    B *= g["spec_mask"][None, :]          # a G2 / G4 input: zero on the masked frequencies
    B64 = B.astype(np.complex64)
    Bd = _cuda(B64)
    yh, yl, se = split_hankel_fp16(Y)
    Yh, Yl = _cuda(yh), _cuda(yl)
    rs = rng.uniform(0.5, 1.5, Nr)
    rsd = _cuda(rs.astype(np.float32))
    slots = torch.zeros(4, dtype=torch.int32, device="cuda")
    sp = lambda i: ctypes.c_void_p(slots.data_ptr() + 4 * i)  # noqa: E731
    st = _lib.stream_ptr()
    _lib.call("fp_absmax_f32", _lib.ptr(Bd), 2 * Nr * Nt, sp(0), st)
    full = torch.full_like(Bd, complex(7.0, 7.0))
    band = torch.full_like(Bd, complex(7.0, 7.0))
    _lib.call("fp_hankel_gemm_mixed", _lib.ptr(Yh), _lib.ptr(Yl), se, _lib.ptr(Bd), _lib.ptr(full),
              Nr, Nt, ctypes.c_float(0.5), _lib.ptr(rsd), None, sp(0), sp(1), None, st)
    _lib.call("fp_hankel_gemm_mixed_band", _lib.ptr(Yh), _lib.ptr(Yl), se, _lib.ptr(Bd),
              _lib.ptr(band), Nr, Nt, ctypes.c_float(0.5), _lib.ptr(rsd), None, sp(0), sp(2),
              None, zlo, zhi, 1, st)
    torch.cuda.synchronize()
    ref = 0.5 * rs.astype(np.float32)[:, None] * (Y @ B64.astype(np.complex128))
    scale = np.abs(ref).max()
    got_b = band.cpu().numpy().astype(np.complex128)
    got_f = full.cpu().numpy().astype(np.complex128)
    e_b = np.abs(got_b - ref).max() / scale
    e_f = np.abs(got_f - ref).max() / scale
    e_bf = np.abs(got_b - got_f).max() / scale
    _announce(f"  vs FP64: band {e_b:.2e} full {e_f:.2e}; band vs full {e_bf:.2e}")
    assert e_b < 4e-6 and e_f < 4e-6 and e_bf < 4e-6
    b = torch.view_as_real(band).reshape(Nr, 2 * Nt)
    live = torch.ones(2 * Nt, dtype=torch.bool, device="cuda")
    live[zlo:zhi] = False
    assert bool((b[:, ~live] == 0).all())            # zero-filled masked columns
    mb = slots[2:3].cpu().numpy().view(np.float32)[0]
    mf = slots[1:2].cpu().numpy().view(np.float32)[0]
    assert abs(mb - mf) <= 4e-6 * mf and mb == float(b.abs().max())
