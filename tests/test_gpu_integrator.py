"""Whole-loop parity: the GPU Integrator (FP64 pipeline) against the CPU oracle
on the same seeded inputs. Tolerance: rel-L2 <= 1e-9 on the field, fluence and
electron density (FP64 on both sides; differences are summation order and
1-ulp libm differences amplified by the reference's phi1 formula)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import reference_numpy as O  # noqa: E402
from tests._cases import C1, C2, cfg_kwargs, envelope_args, rel_l2  # noqa: E402


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _pair(case, steps=None, **over):
    import rapport_amd as R
    kw = cfg_kwargs(case, **over)
    env = envelope_args(case)
    it_p = R.Integrator(R.Config(**kw), envelope=R.envelope_lens_noise(**env))
    it_o = O.Integrator(O.Config(**kw), envelope=O.envelope_lens_noise(**env))
    res_p = it_p.propagate(steps)
    res_o = it_o.propagate(steps)
    return it_p, it_o, res_p, res_o


def _compare(res_p, res_o, keys, tol):
    for k in keys:
        a, b = res_p[k], res_o[k]
        if b is None:
            assert a is None, k
            continue
        assert a.shape == b.shape, (k, a.shape, b.shape)
        if np.abs(b).max() == 0:
            assert np.abs(a).max() == 0, k
            continue
        e = rel_l2(a, b)
        assert e <= tol, (k, e)


KEYS = ["z", "fluence_rz", "rho_rz", "rho_s_rz", "Imax_z", "absorbed_rz", "Ipeak_rz",
        "E_MPI_z", "E_plasma_z", "E_total_z", "rho_onaxis_t", "I_onaxis_t", "r", "rlist",
        "t_full_fs"]


def test_c1_full_run_vs_oracle():
    _announce("integrator C1 200 steps")
    it_p, it_o, res_p, res_o = _pair(C1, save_stride=10)
    assert len(res_p["z"]) == 20
    _compare(res_p, res_o, KEYS, 1e-9)
    assert rel_l2(it_p.u.cpu().numpy(), it_o.u) < 1e-9


def test_hot_run_all_channels_vs_oracle():
    # tight beam: ionisation, plasma, Raman, STE index, absorber bins, snapshot and cube
    _announce("integrator hot 40 steps, STE on")
    hot = dict(C1, nz=39, end=39 * 10e-6, energy_uJ=20.0)
    it_p, it_o, res_p, res_o = _pair(hot, w0=30e-6, R_factor=20.0, save_stride=4,
                                     enable_ste=True, tau_ste=1e-12, rho_t_stride=7,
                                     rho_r_stride=3, absorb_time_bins_fs=(-100, -50, 0, 50, 100),
                                     rho_snapshot_t_fs=40.0)
    assert res_o["rho_rz"].max() > 1e17
    _compare(res_p, res_o, KEYS + ["rho_rzt", "rho_s_rzt", "I_rzt", "t_sub_fs", "r_sub",
                                   "absorbed_rz_bins", "rho_rz_at", "rho_s_onaxis_t"], 1e-9)
    assert rel_l2(it_p.u.cpu().numpy(), it_o.u) < 1e-9


def test_c2_two_steps_vs_oracle():
    _announce("integrator C2 shape 2 steps")
    two = dict(C2, nz=1, end=5e-6)
    it_p, it_o, res_p, res_o = _pair(two, save_stride=1)
    _compare(res_p, res_o, ["fluence_rz", "rho_rz", "Imax_z", "Ipeak_rz"], 1e-9)
    assert rel_l2(it_p.u.cpu().numpy(), it_o.u) < 1e-10


def test_step_api_matches_pipeline():
    # Integrator.step (time-domain, reference API) == propagate()'s spectral pipeline
    import rapport_amd as R
    _announce("step API vs pipeline")
    kw = cfg_kwargs(C1, energy_uJ=40.0, w0=40e-6, R_factor=25.0)
    env = R.envelope_lens_noise(**envelope_args(C1))
    a = R.Integrator(R.Config(**kw), envelope=env)
    b = R.Integrator(R.Config(**kw), envelope=env)
    for i in range(3):
        a.nl.update_plasma(a.u, a.rho, a.rho_s, a.dt)
        a.step(a.dz)
        b._pipeline_step(b.dz)
    # same arithmetic up to FFT placement; 3 hot nonlinear steps amplify the rounding
    assert rel_l2(a.u.cpu().numpy(), b.u.cpu().numpy()) < 1e-10
    assert rel_l2(a.rho.cpu().numpy(), b.rho.cpu().numpy()) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_matches_eager(precision):
    # a captured z-step replayed 4 times, with eager work on the stream between
    # replays, gives bit-identical state to 4 eager steps
    import rapport_amd as R
    _announce(f"graph replay vs eager, {precision}")
    kw = cfg_kwargs(C1, energy_uJ=20.0, w0=30e-6, R_factor=20.0, precision=precision)
    env = R.envelope_lens_noise(**envelope_args(C1))
    a = R.Integrator(R.Config(**kw), envelope=env)
    b = R.Integrator(R.Config(**kw), envelope=env)
    g = b.capture_step(1)
    junk = torch.zeros(1 << 20, device="cuda")
    for i in range(4):
        a._pipeline_step(a.dz)
        g.replay()
        junk.add_(1.0)
        torch.cuda.synchronize()
    assert torch.equal(a.u, b.u)
    assert torch.equal(a.rho, b.rho)
    assert float(junk[0]) == 4.0


@pytest.mark.parametrize("precision,Nt", [("fp64", 512), ("mixed", 512), ("fp64", 1024),
                                          ("fp64", 2048), ("mixed", 2048), ("mixed", 4096)])
def test_fused_record_matches_unfused(precision, Nt):
    """fp_record_history (inverse transform, reductions and every history entry from the
    spectral state, two launches) against the unfused record: inverse fp_fft_rows, then
    fp_record_rows, then the per-history torch expressions of integrator.py:182-224.
    Both sides are FP64 up to the float32 stores, so the histories agree to float32
    rounding (the per-row sums run in another order when Nt < 2048, and at Nt = 4096 the row
    kernel's transform runs on 512 threads against 256 in fp_fft_rows). The sizes cover the
    four block shapes of the row kernel (64, 128, 256 and 512 threads)."""
    import ctypes

    import rapport_amd as R
    from rapport_amd._lib import call, ptr, stream_ptr
    _announce(f"fused record vs unfused, {precision} Nt={Nt}")
    hot = dict(C1, Nt=Nt, nz=11, end=11 * 10e-6, energy_uJ=20.0)
    kw = cfg_kwargs(hot, w0=30e-6, R_factor=20.0, save_stride=4, enable_ste=True, tau_ste=1e-12,
                    rho_t_stride=7, rho_r_stride=3, absorb_time_bins_fs=(-100, -50, 0, 50, 100),
                    rho_snapshot_t_fs=40.0, precision=precision)
    it = R.Integrator(R.Config(**kw), envelope=R.envelope_lens_noise(**envelope_args(hot)))
    it.propagate()
    assert float(it.rho.max()) > 1e17

    u = it.u
    nb = int(it._bin_lo.numel())
    rows = torch.zeros(7, it.Nr, dtype=torch.float64, device=u.device)
    bins = torch.zeros(nb, it.Nr, dtype=torch.float64, device=u.device)
    call("fp_record_rows", ptr(u), ptr(it.rho), ptr(it.rho_s), ptr(rows),
         ctypes.byref(it.nl.params(it.dz)), ctypes.byref(it.t.tables), it.dt, nb,
         ptr(it._bin_lo), ptr(it._bin_hi), ptr(bins), stream_ptr())
    invE2, dt, rs, ts = it.g["invE2"], it.dt, it._r_stride, it._t_stride
    I = (u.real ** 2 + u.imag ** 2) * invE2
    expect = {
        "fluence_rz": (rows[0] * dt * invE2).float(), "rho_rz": rows[1].float(),
        "rho_s_rz": rows[2].float(), "Ipeak_rz": rows[3].float(), "absorbed_rz": rows[4].float(),
        "absorbed_rz_bins": bins.float(), "rho_rz_at": it.rho[:, it._it_snap].float(),
        "rho_rzt": it.rho[::rs, ::ts].float(), "rho_s_rzt": it.rho_s[::rs, ::ts].float(),
        "I_rzt": I[::rs, ::ts].float(), "rho_onaxis_t": it.rho[0].float(),
        "rho_s_onaxis_t": it.rho_s[0].float(), "I_onaxis_t": I[0].float(),
    }
    imax = float(rows[3].max())
    dedz = [float((rows[q] * dt * it._r_w_dev).sum() * 1e4 * 1e6) for q in (5, 6)]

    k = 1                                          # overwrite record 1 with the current state
    others = {name: (getattr(it, name)[k - 1].clone(), getattr(it, name)[k + 1].clone())
              for name in expect}
    for name in expect:
        getattr(it, name)[k].fill_(-1.0)
    it.k_save = k
    it._record(0.0)
    torch.cuda.synchronize()
    assert it.k_save == k + 1
    for name, e in expect.items():
        got = getattr(it, name)[k]
        assert got.shape == e.shape, name
        # pointwise samples: transform rounding is relative to the row's peak, so the
        # absolute tolerance is 1e-6 of the array's maximum
        torch.testing.assert_close(got, e, rtol=1e-6, atol=1e-6 * float(e.abs().max()), msg=name)
    # the FP64 scalars: exact / summation-order close where the field is bitwise the unfused
    # one; FP32 transform rounding where the row kernel's transform differs (mixed, 4096)
    same_field = precision == "fp64" or Nt < 4096
    if same_field:
        assert float(it._Imax_dev[k]) == imax
    np.testing.assert_allclose(float(it._Imax_dev[k]), imax, rtol=1e-6)
    rt = 1e-12 if same_field else 1e-6
    np.testing.assert_allclose(float(it._dEdz_dev[0, k]), dedz[0], rtol=rt)
    np.testing.assert_allclose(float(it._dEdz_dev[1, k]), dedz[1], rtol=rt)
    assert dedz[0] > 0 and dedz[1] > 0
    for name, (before, after) in others.items():   # the neighbouring records are untouched
        assert torch.equal(getattr(it, name)[k - 1], before), name
        assert torch.equal(getattr(it, name)[k + 1], after), name


# params.json fields written by the reference (D3/sim/integrator.py:305-340)
PARAMS_KEYS = {"n0", "n0_probe", "nc_probe_cm3", "lambda_probe_nm", "wavelength_nm", "n2", "U_g_eV",
               "Us_eV", "energy_uJ", "w0_um", "delta_t_fs", "code_fingerprint", "begin_um", "end_um",
               "z_focus_air_um", "toggles"}
TOGGLE_KEYS = {"enable_kerr_instantaneous", "enable_kerr_raman", "enable_self_steepening",
               "enable_photoionization_loss", "enable_plasma_defocusing", "enable_plasma_absorption",
               "enable_space_time_focusing", "enable_spectral_filter", "tau_ste_fs", "enable_ste_index",
               "E_tr_eV", "f_ste_pump", "enable_avalanche", "enable_recombination", "enable_ste"}


def test_run_writes_reference_outputs(tmp_path):
    """run() through the reference's entry point with its defaults (recording every step,
    cube subsampling, STE on): result.npz holds the same keys and values as the oracle's,
    params.json the reference's fields."""
    import json
    import rapport_amd as R
    _announce("run() writes result.npz / params.json")
    kw = dict(Nz=6, Nt=512, Nr=129, wavelength=800e-9, energy_uJ=20.0, w0=30e-6, delta_t=85e-15,
              begin=0.0, end=60e-6, n2=3.2e-20, R_factor=20.0, verbose=False, t_window=800e-15,
              rho_t_stride=15)
    res_p = R.run(out_dir=str(tmp_path / "gpu"), **kw)
    res_o = O.run(out_dir=str(tmp_path / "cpu"), **kw)
    zp = np.load(tmp_path / "gpu" / "result.npz", allow_pickle=True)      # written by this test
    zo = np.load(tmp_path / "cpu" / "result.npz", allow_pickle=True)
    assert set(zp.files) == set(zo.files) == set(res_o)
    for k in zo.files:
        a, b = zp[k], zo[k]
        if b.dtype == object:                       # None entries (disabled outputs)
            assert a.dtype == object and a.item() is None and b.item() is None, k
            continue
        assert a.shape == b.shape, k
        if np.abs(b).max() == 0:
            assert np.abs(a).max() == 0, k
        else:
            assert rel_l2(a, b) <= 1e-9, (k, rel_l2(a, b))
        np.testing.assert_array_equal(a, res_p[k] if res_p[k] is not None else a)
    assert res_o["rho_rzt"] is not None and res_o["rho_s_rzt"] is not None   # the cube is on
    params = json.load(open(tmp_path / "gpu" / "params.json"))
    assert set(params) == PARAMS_KEYS and set(params["toggles"]) == TOGGLE_KEYS
    assert params["toggles"]["enable_ste"] is True
    assert abs(params["n0"] - json.load(open(tmp_path / "cpu" / "params.json"))["n0"]) == 0.0


def test_in_place_edit_of_u_sets_the_field():
    # the reference's Integrator.u is an array (integrator.py:40): an in-place write into it
    # is the field the next step starts from. Same here for the tensor `u` returns.
    import rapport_amd as R
    _announce("in-place edit of Integrator.u")
    kw = cfg_kwargs(C1, nz=4, end=4 * 10e-6)
    env = envelope_args(C1)
    for prec in ("fp64", "mixed"):
        a = R.Integrator(R.Config(**kw, precision=prec), envelope=R.envelope_lens_noise(**env))
        b = R.Integrator(R.Config(**kw, precision=prec), envelope=R.envelope_lens_noise(**env))
        ua = a.u
        ua[:, : a.Nt // 2] *= 0.5                    # in place
        ub = b.u
        ub[:, : b.Nt // 2] *= 0.5
        b.u = ub                                     # explicit set
        assert torch.equal(a.u, b.u)
        ra, rb = a.propagate(), b.propagate()
        assert np.array_equal(ra["fluence_rz"], rb["fluence_rz"]), prec
        assert torch.equal(a.u, b.u), prec
        untouched = R.Integrator(R.Config(**kw, precision=prec), envelope=R.envelope_lens_noise(**env))
        assert not np.array_equal(untouched.propagate()["fluence_rz"], ra["fluence_rz"])


def _scenarios():
    """The reference's ablation scenarios (term_ablation_study.ipynb:1242-1250): full,
    one no_<term> per field toggle, and linear_only (all six off)."""
    from rapport_amd import FIELD_TOGGLES
    sc = {"full": {}}
    for f in FIELD_TOGGLES:
        sc["no_" + f.replace("enable_", "")] = {f: False}
    sc["linear_only"] = {f: False for f in FIELD_TOGGLES}
    return sc


# 12 uJ into 30 um: every scenario ionises (rho up to 1e17-1e18 cm^-3) and stays finite; at
# 20 uJ the reference arithmetic itself overflows without plasma absorption (dz = 10 um)
HOT = dict(C1, nz=19, end=19 * 10e-6, energy_uJ=12.0)
HOT_KW = dict(w0=30e-6, R_factor=20.0, save_stride=2, enable_ste=True, tau_ste=1e-12)


@pytest.mark.parametrize("name", list(_scenarios()))
def test_ablation_scenario_vs_oracle(name):
    # each ablation scenario, 20 hot steps (ionisation, plasma, Raman, STE), FP64 against
    # the oracle at 1e-9 and the mixed path against the oracle at its 1e-4 fluence bar
    import rapport_amd as R
    over = _scenarios()[name]
    _announce(f"ablation scenario {name}: {over}")
    it_p, it_o, res_p, res_o = _pair(HOT, **HOT_KW, **over)
    assert res_o["rho_rz"].max() > 1e17 and np.isfinite(res_o["fluence_rz"]).all()
    _compare(res_p, res_o, KEYS, 1e-9)
    assert rel_l2(it_p.u.cpu().numpy(), it_o.u) < 1e-9
    kw = cfg_kwargs(HOT, **HOT_KW, **over, precision="mixed")
    res_m = R.Integrator(R.Config(**kw), envelope=R.envelope_lens_noise(**envelope_args(HOT))
                         ).propagate()
    e = rel_l2(res_m["fluence_rz"], res_o["fluence_rz"])
    _announce(f"  mixed fluence rel-L2 {e:.2e}")
    assert e <= 1e-4


@pytest.mark.parametrize("over", [dict(keldysh_monotone=True),
                                  dict(keldysh_beta_den=2.0),
                                  dict(z_focus_air_um=150.0)])
def test_host_setup_variants_vs_oracle(over):
    # host-setup variants: the monotone Keldysh envelope and beta_den = 2 (keldysh.py:80-148)
    # behind the rate tables, and begin from z_focus_air_um (config.py:210-215: begin =
    # -n0 z_focus_air; end is moved with it so that dz stays 10 um), through a hot FP64 run
    # against the oracle
    _announce(f"host-setup variant {over}")
    over = dict(over)
    if "z_focus_air_um" in over:
        over["end"] = -O.n_sellmeier(800e-9) * over["z_focus_air_um"] * 1e-6 + HOT["end"]
    it_p, it_o, res_p, res_o = _pair(HOT, **HOT_KW, **over)
    assert res_o["rho_rz"].max() > 1e17
    _compare(res_p, res_o, KEYS, 1e-9)
    if "z_focus_air_um" in over:
        assert it_p.cfg.begin == it_o.cfg.begin == -it_o.cfg.n0 * 150.0e-6
        assert abs(it_o.cfg.dz - 10e-6) < 1e-12
        assert np.allclose(res_p["z"], res_o["z"]) and res_p["z"][0] < 0
