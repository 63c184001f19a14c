"""CPU tests of the product's host side: the C-ABI library loads and exports
every declared entry point, and the host setup (Config, Keldysh tables, grids,
envelopes) equals the oracle's."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import rapport_amd as R
from oracle import reference_numpy as O
from tests._cases import C1, C2, cfg_kwargs, envelope_args

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "rapport_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(fp_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from rapport_amd import _lib
    lib = _lib.load()
    decl = _declared()
    assert decl == _lib.declared_symbols()
    for name in decl:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (fp_\w+)", out))
    assert set(decl) <= exported
    assert lib.fp_version() >= 100


def test_gpu_entry_points_refuse_cpu_tensors():
    import torch
    it_cfg = R.Config(**cfg_kwargs(C1))
    g = R.build_grids(it_cfg)
    from rapport_amd import _lib
    with pytest.raises(ValueError):
        _lib.ptr(torch.zeros(3))


@pytest.mark.parametrize("case", [C1, C2])
def test_config_and_grids_match_oracle(case):
    kw = cfg_kwargs(case)
    cp, co = R.Config(**kw), O.Config(**kw)
    for a in ("omega0", "n0", "chi3", "komega", "dz", "b", "E0", "I0_Wcm2", "Omega_r2", "nc_probe"):
        assert np.isclose(getattr(cp, a), getattr(co, a), rtol=1e-15, atol=0), a
    gp, go = R.build_grids(cp), O.build_grids(co)
    for k in ("tmax", "dt", "tlist", "ff", "Y", "R", "j1last", "rlist", "rholist", "delta_k", "k1",
              "T_op", "inv_U_op", "inv_U_nl", "R_f", "spec_mask", "rr", "tt", "rhorho",
              "sigmaomega", "avalanche_coef", "avalanche_coef_s", "inv_taur_eff", "inv_tau_ste",
              "ste_pref", "f_ste", "invE2", "mask_r", "b", "E0", "komega"):
        a, b = np.asarray(gp[k]), np.asarray(go[k])
        assert a.shape == b.shape, k
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-13 * scale, k
    kp, ko = gp["keldysh"], go["keldysh"]
    for k in ("W_LUT", "W_LUT_s", "NLUT", "logImin", "inv_dlog", "Imin", "Imax"):
        assert np.allclose(kp[k], ko[k], rtol=1e-14, atol=0), k
    assert np.array_equal(kp["f_spline"].x, ko["f_spline"].x)
    assert np.allclose(kp["f_spline"].c, ko["f_spline"].c, rtol=1e-14, atol=0)


def test_envelopes_match_oracle():
    kw = cfg_kwargs(C2)
    cp, co = R.Config(**kw), O.Config(**kw)
    gp, go = R.build_grids(cp), O.build_grids(co)
    up = R.envelope_lens_noise(**envelope_args(C2))(gp["rr"], gp["tt"], cp, gp)
    uo = O.envelope_lens_noise(**envelope_args(C2))(go["rr"], go["tt"], co, go)
    assert np.abs(up - uo).max() <= 1e-14 * np.abs(uo).max()
    u1 = R.envelope_gaussian_focused(gp["rr"], gp["tt"], cp, gp)
    r0 = gp["rlist"][0]
    assert np.abs(u1[0]).max() == pytest.approx(cp.E0 * np.exp(-r0 ** 2 / cp.w0 ** 2), rel=1e-12)


def test_keldysh_validate_product():
    from rapport_amd import keldysh
    v = keldysh.validate(verbose=False)
    assert f"{v['W_3.5e13_beta4']:.2e}" == "6.44e+32"
    assert f"{v['W_3.5e13_beta2']:.2e}" == "4.30e+32"
    assert np.allclose(keldysh.keldysh_multiphoton(np.array([1e11])),
                       O.Keldysh(800e-9, 9.0, 0.64 * 9.1093837015e-31, O.n_sellmeier(800e-9)).rate(np.array([1e11])),
                       rtol=0.3)


def test_fingerprint_changes_with_sources(tmp_path):
    fp = R.code_fingerprint()
    assert re.fullmatch(r"[0-9a-f]{12}", fp)


def _record_args(Nr=128, Nt=512, nbins=0, **dst):
    """Arguments of fp_record_history with placeholder device addresses (never
    dereferenced: every case below is rejected before a launch)."""
    import ctypes
    from rapport_amd import _lib
    p = _lib.NLParams()
    p.Nr, p.Nt = Nr, Nt
    t = _lib.NLTables()
    d = _lib.RecordDst()
    for f in ("fluence", "rho", "rho_s", "ipeak", "absorbed", "rho_onaxis", "rho_s_onaxis",
              "I_onaxis", "r_w", "partial", "imax", "dedz_photo", "dedz_plasma"):
        setattr(d, f, 0x1000)
    d.r_stride, d.t_stride, d.Nt_sub = 1, 0, 0
    for k, v in dst.items():
        setattr(d, k, v)
    lohi = 0x1000 if nbins else None
    return [0x1000, 0, 0x1000, 0x1000, None, ctypes.byref(p), ctypes.byref(t), 1e-16, nbins,
            lohi, lohi, ctypes.byref(d), None]


@pytest.mark.parametrize("kw,msg", [
    (dict(fluence=None), "null destination"),
    (dict(imax=None), "null destination"),
    (dict(nbins=2), "bins"),
    (dict(rho_at=0x1000, it_snap=512), "it_snap"),
    (dict(rho_at=0x1000, it_snap=-1), "it_snap"),
    (dict(t_stride=-1), "t_stride"),
    (dict(t_stride=7, r_stride=0, Nt_sub=74), "r_stride"),
    (dict(t_stride=7, r_stride=3, Nt_sub=73), "Nt_sub"),
    (dict(t_stride=7, r_stride=3, Nt_sub=74), "null cube"),
    (dict(Nt=300), "Nt must be"),
    (dict(Nr=0), "Nr must be positive"),
])
def test_record_history_rejects_bad_arguments(kw, msg):
    """fp_record_history validates its destinations and shapes on the host: the histories
    are written at raw addresses, so a wrong stride or index must not reach a kernel."""
    from rapport_amd import _lib
    shape = {k: kw.pop(k) for k in ("Nr", "Nt", "nbins") if k in kw}
    args = _record_args(**shape, **kw)
    with pytest.raises(RuntimeError, match=msg):
        _lib.call("fp_record_history", *args)


@pytest.mark.parametrize("kw,msg", [
    (dict(Nr=16385), "FP_NOALLOC_MAX_NR"),
    (dict(Nr=0), "Nr must be positive"),
    (dict(fluence=None), "null destination"),
    (dict(t_stride=7, r_stride=3, Nt_sub=73), "Nt_sub"),
])
def test_record_history_noalloc_rejects_bad_arguments(kw, msg):
    """The allocation-free record has the same host checks, plus its row limit (the
    static scratch holds 3 x 16384 partials); none of these reaches the GPU."""
    from rapport_amd import _lib
    shape = {k: kw.pop(k) for k in ("Nr", "Nt", "nbins") if k in kw}
    args = _record_args(**shape, **kw)
    with pytest.raises(RuntimeError, match=msg):
        _lib.call("fp_record_history_noalloc", *args)


def test_record_history_rejects_null_spec():
    from rapport_amd import _lib
    args = _record_args()
    args[0] = None
    with pytest.raises(RuntimeError, match="null pointer"):
        _lib.call("fp_record_history", *args)


@pytest.mark.parametrize("Nr,Nt,msg", [
    (200, 512, "Nr % 128"), (0, 512, "Nr % 128"), (128, 100, "2 Nt % 128"),
])
def test_mixed_gemm_rejects_ragged_shapes(Nr, Nt, msg):
    """The split-FP16 GEMM entry point tiles the radius by 128: a ragged operand is refused
    on the host before any launch (the mixed engine pads its operands, see below)."""
    from rapport_amd import _lib
    from rapport_amd.mixed import split_hankel_fp16
    a = 0x1000
    with pytest.raises(RuntimeError, match=msg):
        _lib.call("fp_hankel_gemm_mixed", a, a, 0, a, 0x2000, Nr, Nt, ctypes.c_float(1.0),
                  None, None, a, None, None, None)
    with pytest.raises(RuntimeError, match="alias"):
        _lib.call("fp_hankel_gemm_mixed", a, a, 0, a, a, 128, 512, ctypes.c_float(1.0),
                  None, None, a, None, None, None)


@pytest.mark.parametrize("Nr", [128, 200, 300])
def test_split_hankel_pads_to_the_tile_height(Nr):
    """split_hankel_fp16 zero-pads Y to Np = 128 ceil(Nr / 128): un-fragmented, hi + lo is
    Y 2^s to FP16-pair precision in the top-left Nr x Nr block and exactly 0 elsewhere."""
    from rapport_amd.mixed import padded_rows, split_hankel_fp16
    rng = np.random.default_rng(Nr)
    Y = rng.standard_normal((Nr, Nr)) * 1e-3
    yh, yl, se = split_hankel_fp16(Y)
    Np = padded_rows(Nr)
    assert Np % 128 == 0 and Nr <= Np < Nr + 128

    def unfrag(F):
        return F.transpose(1, 3, 0, 2, 4).reshape(Np, Np).astype(np.float64)
    full = (unfrag(yh) + unfrag(yl)) * 2.0 ** -se
    assert np.abs(full[:Nr, :Nr] - Y).max() <= 2.0 ** -21 * np.abs(Y).max()
    assert not full[Nr:].any() and not full[:, Nr:].any()
    with pytest.raises(ValueError, match="square"):
        split_hankel_fp16(np.zeros((Nr, Nr + 1)))


def test_header_is_a_c_abi(tmp_path):
    """include/rapport_amd.h compiles as C99, and a C program links the library through it
    and calls an entry point (fp_version needs no GPU)."""
    import shutil
    from rapport_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "abi.c"
    src.write_text('#include "rapport_amd.h"\n#include <stdio.h>\n'
                   "int main(void) {\n"
                   "  fp_record_dst d = {0}; fp_nl_params p = {0}; (void)d; (void)p;\n"
                   '  printf("%d\\n", fp_version());\n'
                   "  return 0;\n}\n")
    exe = tmp_path / "abi"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-L", libdir, "-lrapport_amd", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True,
                         env={**os.environ, "LD_LIBRARY_PATH": libdir})
    assert int(out.stdout) >= 100


def test_fp16_scale_helpers_at_extremes(tmp_path):
    """pow2f (the GEMM's power-of-two scales and the epilogue's unscale 2^-(sY + sB)) is
    exact over the whole float range, subnormals included, 0 below 2^-149 and inf above
    2^127; scale_exp keeps max|B| 2^se in [2^12, 2^13) (FP16 headroom) and clamps to
    [-100, 100], returning 0 for 0, negative, inf and NaN maxima. Compiled host-only."""
    import math
    import shutil
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "scale_check"
    src = os.path.join(ROOT, "tests", "host_checks", "scale_check.cpp")
    r = subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1",
                        "-I", os.path.join(ROOT, "rapport_amd", "csrc"), src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    p2, se = {}, []
    for ln in out.splitlines():
        k, a, b = ln.split()
        if k == "p":
            p2[int(a)] = float.fromhex(b)
        else:
            se.append((float.fromhex(a) if "0x" in a else float(a), int(b)))
    for e, v in p2.items():
        ref = math.inf if e > 127 else (math.ldexp(1.0, e) if e >= -149 else 0.0)
        assert v == ref, (e, v, ref)
    for x, e in se:
        if not (x > 0) or math.isinf(x) or math.isnan(x):
            assert e == 0, (x, e)
            continue
        assert -100 <= e <= 100, (x, e)
        if -100 < e < 100:
            assert 2.0 ** 12 <= x * 2.0 ** e < 2.0 ** 13, (x, e)
        elif e == 100:
            assert x * 2.0 ** 100 < 2.0 ** 13
        else:
            assert x * 2.0 ** -100 >= 2.0 ** 12


@pytest.mark.parametrize("name", ["fp_nl_split", "fp_nl_step", "fp_record_rows",
                                  "fp_nl_split_c64", "fp_nl_step_c64"])
def test_nonlinear_entry_points_reject_nonpositive_rows(name):
    """A row count <= 0 is a shape error on the host (FP_E_SHAPE), not a launch error."""
    from rapport_amd import _lib
    fake = ctypes.c_void_p(0x1000)
    p = _lib.NLParams(Nr=0, Nt=512)
    t32 = _lib.NLTablesF32(*([fake] * 9))
    t64 = _lib.NLTables(*([fake] * 8))
    args = {
        "fp_nl_split": (fake, fake, None, fake, fake, ctypes.byref(p), ctypes.byref(t64), None),
        "fp_nl_step": (fake, fake, fake, None, 1, 1, ctypes.byref(p), ctypes.byref(t64), None),
        "fp_record_rows": (fake, fake, None, fake, ctypes.byref(p), ctypes.byref(t64), 1e-15, 0,
                           None, None, None, None),
        "fp_nl_split_c64": (fake, fake, None, fake, fake, ctypes.byref(p), ctypes.byref(t32),
                            None),
        "fp_nl_step_c64": (fake, fake, fake, None, 1, 1, ctypes.byref(p), ctypes.byref(t32), None,
                           None),
    }[name]
    with pytest.raises(RuntimeError, match="Nr must be positive"):
        _lib.call(name, *args)


def test_exact_twiddle_table_layout(tmp_path):
    """The compact twiddle tables that mixed.exact_twiddles appends (per pass, r, k: w^r
    rounded and its residual) sit where the exact transform reads them: the table sizes and
    every pass offset of fft.hpp (compiled host-only) match the host layout, and each entry
    is exp(-2 pi i k r / (Ns R)) split into a complex64 part and its residual."""
    import shutil
    from rapport_amd.mixed import _exact_pass_table, exact_twiddles
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "exact_table_check"
    src = os.path.join(ROOT, "tests", "host_checks", "exact_table_check.cpp")
    r = subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1",
                        "-I", os.path.join(ROOT, "rapport_amd", "csrc"), src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    sizes, offs = {}, []
    for ln in out.splitlines():
        f = ln.split()
        if f[0] == "t":
            sizes[int(f[1])] = (int(f[2]), int(f[3]), int(f[4]))
        else:
            offs.append(tuple(int(x) for x in f[1:]))
    for N, (n8, n4, n16) in sizes.items():
        assert n8 == len(_exact_pass_table(N, 8)) and n4 == len(_exact_pass_table(N, 4)), N
        assert n16 == len(_exact_pass_table(N, 16)), N
        assert exact_twiddles(N).size == 2 * N + 2 * (n8 + n4 + n16)
    tw = {N: exact_twiddles(N).view(np.float32) for N in sizes}
    for N, E, Ns, R, off in offs:
        base = 4 * N + 4 * off           # offsets include the tables stored before E's
        tab = tw[N][base: base + 4 * Ns * (R - 1)].reshape(R - 1, Ns, 4).astype(np.float64)
        k = np.arange(Ns)[None, :]
        rr = np.arange(1, R)[:, None]
        z = np.exp(-2j * np.pi * k * rr / (Ns * R))
        assert np.abs(tab[..., 0] + tab[..., 2] - z.real).max() < 1e-14
        assert np.abs(tab[..., 1] + tab[..., 3] - z.imag).max() < 1e-14
        assert np.abs(tab[..., 0] - z.real).max() <= 2 ** -24
    assert len(offs) > 0


def test_lds_swizzle_bijective_linear_and_conflict_free(tmp_path):
    """The row transforms' LDS swizzle (fft.hpp lds_pad, compiled host-only): a bijection on
    [0, N), linear over XOR, and the base-XOR-constant write and base-plus-offset read
    addresses the transforms use equal lds_pad of the element index, for every row shape
    (N = 512 .. 4096, 4, 8 or 16 samples per thread). Under a 32-bank x 4-byte LDS model (16
    lanes of 8-byte accesses per cycle) the Stockham writes and stride-layout reads of the
    complex64 transforms at 8 and 16 samples per thread, and of the complex128 ones at 8, are
    free of bank conflicts, and every stride-layout read is. The model is the one the GPU's
    counters bear out: without bit 7 -> bit 0 in the swizzle the 16-sample complex64
    transform measured 1.3 conflict cycles per LDS instruction and the 8-sample ones none,
    which this model predicts (1 and 0) and a 64-bank one does not."""
    import shutil
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "lds_swizzle_check"
    src = os.path.join(ROOT, "tests", "host_checks", "lds_swizzle_check.cpp")
    r = subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1",
                        "-I", os.path.join(ROOT, "rapport_amd", "csrc"), src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    rows, fails = [], None
    for ln in out.splitlines():
        f = ln.split()
        if f[0] == "f":
            fails = int(f[1])
        else:
            rows.append(tuple(int(x) for x in f[1:]))
    assert fails == 0
    assert len(rows) == 22
    for N, tpb, es, wdeg, rdeg in rows:
        assert rdeg == 1, (N, tpb, es)
        if (es == 8 and N // tpb in (8, 16)) or (es == 16 and N // tpb == 8):
            assert wdeg == 1, (N, tpb, es, wdeg)


def test_pchip_interval_estimate_is_one_fixup_from_searchsorted():
    # the FP64 nonlinear step finds the PCHIP interval without loops: decade from the binary
    # exponent ((e2 * 1233) >> 12, one step up against 10^(k+1)), index from the scaled
    # mantissa, then one comparison each way with the knots. That is exact only if the
    # index before the fix-up is at most one interval off; checked here on the product's
    # knots, every knot and its two neighbouring doubles, and random intensities, with the
    # product's (x * 10^-k - 1) * 100 rounded once (fused) or twice.
    from rapport_amd.grids import intensity_knots
    knots = intensity_knots()
    n = len(knots)
    P10 = 10.0 ** np.arange(18)
    IP10 = 10.0 ** -np.arange(18.0)
    rng = np.random.default_rng(0)
    xs = np.concatenate([knots, np.nextafter(knots, 0), np.nextafter(knots, np.inf),
                         10 ** rng.uniform(0, 17, 200_000)])
    xs = xs[(xs >= knots[0]) & (xs <= knots[-1])]
    e2 = ((xs.view(np.int64) >> 52) & 0x7ff) - 1023
    k = np.minimum(np.where(e2 <= 0, 0, (e2 * 1233) >> 12), 16)
    up = (k < 16) & (xs >= P10[k + 1])
    k = k + up
    true = np.clip(np.searchsorted(knots, xs, side="right") - 1, 0, n - 2)
    for fused in (False, True):
        if fused:
            v = (np.longdouble(xs) * np.longdouble(IP10[k]) - 1).astype(np.float64) * 100.0
        else:
            v = (xs * IP10[k] - 1.0) * 100.0
        i = np.minimum(k * 900 + np.clip(v.astype(np.int64), 0, 899), n - 2)
        assert np.abs(i - true).max() <= 1
        inc = (i < n - 2) & (knots[np.minimum(i + 1, n - 1)] <= xs)
        dec = (i > 0) & (knots[i] > xs)
        assert np.array_equal(np.where(inc, i + 1, np.where(dec, i - 1, i)), true)
