"""Pins for the CPU oracle (oracle/reference_numpy.py).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it: recorded outputs
of the reference (D3 = .github/deliverable (3)), closed forms, invariants and
an independent scalar restatement on tiny inputs. No GPU.
"""
import numpy as np
import pytest
from scipy.constants import c, epsilon_0, m_e
from scipy.constants import elementary_charge as q_e

from oracle import reference_numpy as O


# ---------------------------------------------------------------- Sellmeier
def test_sellmeier_recorded_values():
    # D3/ANALYSE_CRITIQUE.md:58 (1.4533 at 800 nm);
    # D3/notebooks/filament_visualization_1030nm.ipynb:126 (1.45004 at 1030 nm)
    assert round(O.n_sellmeier(800e-9), 4) == 1.4533
    assert round(O.n_sellmeier(1030e-9), 5) == 1.45004


def test_gvd_from_delta_k():
    # k''(800 nm) = 361.6 fs^2/cm (D3/ANALYSE_CRITIQUE.md:59), read off the
    # curvature of build_grids' delta_k around Omega = 0 (delta_k = k - k0 - k1 Omega).
    cfg = O.Config(nz=1, Nt=4096, N=9, wavelength=800e-9, delta_t=85e-15, w0=100e-6,
                   R_factor=10, t_window=20e-12, write_output=False, verbose=False)
    g = O.build_grids(cfg)
    Om = 2 * np.pi * g["ff"]
    sel = np.abs(Om) < 2 * np.pi * 2e12
    p = np.polyfit(Om[sel], g["delta_k"][sel], 3)
    k2 = 2 * p[1]                                # s^2/m
    assert abs(k2 * 1e30 / 100 - 361.6) < 0.5    # fs^2/cm
    assert abs(p[2]) < 1e-6 * abs(g["k1"])       # linear term removed by k1 (central diff)


# ---------------------------------------------------------------- Keldysh
def test_keldysh_recorded_validation_beta_den_2():
    # D3/ANALYSE_CRITIQUE.md:165-175 was recorded with beta_den = 2 (:158)
    n0 = O.n_sellmeier(800e-9)
    k = O.Keldysh(800e-9, 9.0, 0.64 * m_e, n0, beta_den=2.0)
    I = np.array([5e11, 1e12, 2e12])
    ratio = k.rate(I) / (9.6e-70 * I ** 6 * 2.1e22)
    assert np.allclose(np.round(ratio, 3), [1.062, 1.018, 0.911])
    order = np.diff(np.log(k.rate(I))) / np.diff(np.log(I))
    assert round(order.mean(), 2) == 5.89
    II = np.logspace(11.5, 14.2, 600)
    s = np.diff(np.log(np.maximum(k.rate_raw(II), 1e-300))) / np.diff(np.log(II))
    assert int((s < 0).sum()) == 23
    assert round(float(s.min()), 2) == -136.49


def test_keldysh_validate_asserts_and_beta_values():
    v = O.keldysh_validate()
    # keldysh.py:241-245
    assert 0.8 < v["mpi_ratio"][1] < 1.25
    assert 0.95 < v["gamma_at_3.5e13"] < 1.05
    assert abs(v["beta_mpi_limit"] - 1.0) < 1e-3
    assert abs(v["exp_alpha_ratio"] - 1.0) < 1e-3
    # D3/ANALYSE_CRITIQUE.md:172: W(3.5e13) = 6.44e32 (beta_den 4) and 4.30e32 (beta_den 2)
    assert f"{v['W_3.5e13_beta4']:.2e}" == "6.44e+32"
    assert f"{v['W_3.5e13_beta2']:.2e}" == "4.30e+32"
    # W_MPI(3.5e13) = 3.71e34 (D3/ANALYSE_CRITIQUE.md:65)
    assert f"{9.6e-70 * 3.5e13 ** 6 * 2.1e22:.2e}" == "3.71e+34"


def test_keldysh_multiphoton_slope_and_positivity():
    n0 = O.n_sellmeier(800e-9)
    k = O.Keldysh(800e-9, 9.0, 0.64 * m_e, n0)
    I = O.keldysh_knots()
    W = k.rate(I)
    assert np.all(np.isfinite(W)) and np.all(W >= 0)
    # K = ceil(Ui / hbar w0) = 6 at 800 nm: low-intensity log-slope ~ 6
    lo = (I > 1e10) & (I < 1e11)
    sl = np.polyfit(np.log(I[lo]), np.log(W[lo]), 1)[0]
    assert 5.5 < sl < 6.5


# ---------------------------------------------------------------- Config KAT
def _cfg_1030():
    T = 1.0 - ((1.45 - 1.0) / (1.45 + 1.0)) ** 2
    n0 = O.n_sellmeier(1030e-9)
    nz = int((800e-6 + n0 * 272e-6) / 24e-9)
    return O.Config(nz=nz, Nt=4096, N=3001, R_factor=90.0, end=800e-6, z_focus_air_um=272.0,
                    save_stride=100, wavelength=1030e-9, energy_uJ=13.0 * T, w0=3e-6,
                    delta_t=263e-15, n2=2.74e-20, Ui_eV=9.0, meff_rel=0.64, tau_c=1.7e-15,
                    tau_r=330e-15, rho_max=2.1e22, Us_eV=6.0, f_R=0.18, tau_d=32e-15,
                    tau_s=12e-15, enable_ste=True, lambda_probe=490e-9, rho_t_stride=20,
                    rho_r_stride=4, write_output=False, verbose=False)


def test_energy_normalisation_kat_1030nm():
    # filament_visualization_1030nm.ipynb:646 "[init] U_beam(0) = 12.561 uJ",
    # :647 "result.npz ~ 0.99 GB non compresse (86 % = cube (z,r,t))", :127-128 Nz, begin
    cfg = _cfg_1030()
    assert cfg.nz == 49767
    assert f"{cfg.begin * 1e6:+.2f}" == "-394.41"
    it = O.Integrator(cfg)
    assert f"{it.U0_uJ:.3f}" == "12.561"
    assert f"{it.size_estimate_gb:.2f}" == "0.99"
    assert round(100 * it.size_cube_frac) == 86


# ---------------------------------------------------------------- Hankel matrix
@pytest.mark.parametrize("N,tol", [(129, 2.4e-10), (513, 3.9e-12)])
def test_qdht_involution(N, tol):
    cfg = O.Config(nz=1, Nt=8, N=N, wavelength=800e-9, w0=100e-6, R_factor=10,
                   write_output=False, verbose=False)
    Y = O.build_grids(cfg)["Y"]
    err = np.abs(Y @ Y - np.eye(N - 1)).max()
    assert err < 1.3 * tol
    assert np.abs(Y - Y.T).max() > 1e-3       # column J1^2 normalisation: not symmetric


def test_qdht_gaussian_pair():
    # (R^2/j_N) Y exp(-r^2/w^2) = (w^2/2) exp(-rho^2 w^2/4)
    cfg = O.Config(nz=1, Nt=8, N=257, wavelength=800e-9, w0=100e-6, R_factor=10,
                   write_output=False, verbose=False)
    g = O.build_grids(cfg)
    w = 100e-6
    f = np.exp(-g["rlist"] ** 2 / w ** 2)
    F = g["R"] ** 2 / g["j1last"] * g["Y"] @ f
    ref = w ** 2 / 2 * np.exp(-g["rholist"] ** 2 * w ** 2 / 4)
    assert np.abs(F - ref).max() / ref.max() < 1e-14


# ---------------------------------------------------------------- Raman kernel
def test_raman_transfer_closed_form():
    cfg = O.Config(nz=1, Nt=256, N=9, wavelength=800e-9, delta_t=85e-15, w0=100e-6,
                   R_factor=10, t_window=800e-15, write_output=False, verbose=False)
    g = O.build_grids(cfg)
    N, dt = cfg.Nt, g["dt"]
    A = cfg.Omega_r2 * cfg.tau_s
    M = N // 2 - 1                       # lags 1..N/2-1 have t > 0 after ifftshift
    k = np.arange(N)
    out = np.zeros(N, complex)
    for sgn in (+1, -1):
        z = np.exp(-dt / cfg.tau_d + sgn * 1j * dt / cfg.tau_s - 2j * np.pi * k / N)
        out += sgn * z * (1 - z ** M) / (1 - z) / (2j)
    out *= A * dt
    assert np.abs(g["R_f"] - out).max() < 1e-12 * np.abs(out).max()


def test_ste_index_prefactor():
    # Delta n_STE(rho_s = 1e20 cm^-3, 800 nm) = 3.1131e-3 (D3/ANALYSE_CRITIQUE.md:258)
    cfg = O.Config(nz=1, Nt=8, N=9, wavelength=800e-9, w0=100e-6, write_output=False,
                   verbose=False)
    g = O.build_grids(cfg)
    dn = g["ste_pref"] / (cfg.omega0 / c) * 1e20
    assert f"{dn:.4e}" == "3.1131e-03"


# ---------------------------------------------------------------- linear step
def _small_cfg(**kw):
    base = dict(nz=4, Nt=128, N=129, wavelength=800e-9, delta_t=85e-15, w0=100e-6,
                R_factor=10, begin=0.0, end=4e-3, t_window=800e-15, energy_uJ=2.0,
                n2=3.2e-20, tau_c=1.27e-15, tau_r=150e-15, enable_ste=False,
                write_output=False, verbose=False)
    base.update(kw)
    return O.Config(**base)


@pytest.mark.parametrize("j", [0, 5, -7])
def test_half_linear_single_frequency_gaussian_beam(j):
    cfg = _small_cfg(enable_spectral_filter=True)
    g = O.build_grids(cfg)
    lin = O.LinearOperator(cfg, g)
    w0, r, t = cfg.w0, g["rlist"], g["tlist"]
    fj = g["ff"][j]
    u = np.exp(-r ** 2 / w0 ** 2)[:, None] * np.exp(2j * np.pi * fj * t)[None, :]
    dz = 2e-3
    out = lin.half_linear(u, dz)
    # effective wavenumber k0 / inv_U(f_j), advance dz/2, plus exp(i dk dz/2) * mask
    keff = g["komega"] / g["inv_U_op"][j]
    q = 1 + 2j * (dz / 2) / (keff * w0 ** 2)
    ref_r = np.exp(-r ** 2 / (w0 ** 2 * q)) / q
    ref = (ref_r[:, None] * np.exp(2j * np.pi * fj * t)[None, :]
           * np.exp(1j * g["delta_k"][j] * dz / 2) * g["spec_mask"][j])
    assert np.abs(out - ref).max() < 1e-12


# ---------------------------------------------------------------- split()
def _field(cfg, g, seed=1, scale=1.0):
    u = O.envelope_lens_noise(noise=0.05, seed=seed)(g["rr"], g["tt"], cfg, g)
    return u * scale


def test_split_all_toggles_off_is_zero():
    cfg = _small_cfg(enable_kerr_instantaneous=False, enable_kerr_raman=False,
                     enable_self_steepening=False, enable_photoionization_loss=False,
                     enable_plasma_defocusing=False, enable_plasma_absorption=False)
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g)
    rho = np.full(u.shape, 1e19)
    rhs, a = nl.split(u, rho, rho)
    assert np.all(rhs == 0) and np.all(a == 0)


def test_split_kerr_only_is_spm():
    cfg = _small_cfg(enable_kerr_raman=False, enable_self_steepening=False,
                     enable_photoionization_loss=False, enable_plasma_defocusing=False,
                     enable_plasma_absorption=False, enable_space_time_focusing=False,
                     enable_spectral_filter=False)
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g)
    rhs, _ = nl.split(u, np.zeros(u.shape), None)
    kappa = 3 * cfg.chi3 * cfg.omega0 ** 2 / (8 * cfg.komega * c ** 2)
    # kappa |u|^2 = k0 n2 I with I = (n0 c eps0 / 2)|u|^2
    assert abs(kappa - 2 * np.pi / cfg.wavelength * cfg.n2 * 0.5 * cfg.n0 * c * epsilon_0) < 1e-12 * kappa
    ref = 1j * (1 - cfg.f_R) * kappa * np.abs(u) ** 2 * u
    assert np.abs(rhs - ref).max() < 1e-12 * np.abs(ref).max()


# split() one term at a time (D3/sim/operators.py:89-146). With self-steepening, space-time
# focusing and the spectral filter off, T = U^-1 = 1; constant-modulus single-frequency
# fields put every product in one FFT bin, so each spectral operator enters as a number.
_NO_NL = dict(enable_kerr_instantaneous=False, enable_kerr_raman=False,
              enable_photoionization_loss=False, enable_plasma_defocusing=False,
              enable_plasma_absorption=False, enable_self_steepening=False,
              enable_space_time_focusing=False, enable_spectral_filter=False)


def _tone(g, m, I_Wcm2):
    """a exp(2 pi i m n / Nt) on every row, |a|^2 invE2 = I: constant |u|, FFT bin m."""
    Nr, Nt = g["rr"].shape
    a = np.sqrt(I_Wcm2 / g["invE2"])
    return np.broadcast_to(a * np.exp(2j * np.pi * m * np.arange(Nt) / Nt), (Nr, Nt)).copy()


def _kappa(cfg):
    return 3 * cfg.chi3 * cfg.omega0 ** 2 / (8 * cfg.komega * c ** 2)


@pytest.mark.parametrize("kw", [dict(enable_photoionization_loss=True),
                                dict(enable_plasma_absorption=True),
                                dict(enable_photoionization_loss=True, enable_plasma_absorption=True)])
def test_split_losses_cancel_alpha_without_spectral_operators(kw):
    # the losses enter NL as -T FFT(photo u) and -plasma_pref rho u, and rhs adds alpha u:
    # with T = U^-1 = 1 they cancel exactly (a sign error would leave 2 alpha u)
    cfg = _small_cfg(**{**_NO_NL, **kw})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g, scale=30.0)
    rho = np.random.default_rng(4).uniform(0, 1e21, u.shape)
    rhs, a = nl.split(u, rho, None)
    assert a.max() > 0
    assert np.abs(rhs).max() <= 1e-12 * np.abs(a * u).max()


def test_split_plasma_defocusing_is_the_drude_index():
    # defocusing alone: rhs = -i omega0 / (2 n0 c rho_c) * x / (1 + x) * rho u, x = (omega0 tau_c)^2,
    # rho_c = eps0 m omega0^2 / e^2 (the Drude plasma index; rho in m^-3 = 1e6 x cm^-3)
    cfg = _small_cfg(**{**_NO_NL, "enable_plasma_defocusing": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g, scale=10.0)
    rho = np.random.default_rng(5).uniform(0, 1e20, u.shape)
    rhs, a = nl.split(u, rho, None)
    w0 = 2 * np.pi * c / cfg.wavelength
    rho_c = epsilon_0 * cfg.meff_drude_rel * m_e * w0 ** 2 / q_e ** 2
    x = (w0 * cfg.tau_c) ** 2
    ref = -1j * w0 / (2 * cfg.n0 * c * rho_c) * x / (1 + x) * (rho * 1e6) * u
    assert np.all(a == 0)
    assert np.abs(rhs - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("m", [3, -5])
@pytest.mark.parametrize("stf", [False, True])
def test_split_single_frequency_spectral_operators(m, stf):
    # Kerr + photo-ionisation on one tone f_m = m / (Nt dt): the Kerr term carries T^2, the
    # photo term T, both U^-1, and alpha u neither: rhs = [U^-1 (i kappa (1-f_R) |a|^2 T^2
    # - T photo) + photo] u with T = 1 + f_m / f0 and U = 1 + k1 2 pi f_m / k
    cfg = _small_cfg(**{**_NO_NL, "enable_kerr_instantaneous": True,
                        "enable_photoionization_loss": True, "enable_self_steepening": True,
                        "enable_space_time_focusing": stf})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _tone(g, m, 4e13)
    rhs, a = nl.split(u, np.zeros(u.shape), None)
    photo = a[0, 0]
    assert photo > 0 and np.allclose(a, photo, rtol=1e-12, atol=0)
    f_m = m / (cfg.Nt * g["dt"])
    T = 1 + f_m / (c / cfg.wavelength)
    Uinv = 1 / max(1 + g["k1"] * 2 * np.pi * f_m / cfg.komega, 0.1) if stf else 1.0
    a2 = np.abs(u[0, 0]) ** 2
    ref = (Uinv * (1j * _kappa(cfg) * (1 - cfg.f_R) * a2 * T ** 2 - T * photo) + photo) * u
    assert abs(T - 1) > 1e-3 and (not stf or abs(Uinv - 1) > 1e-4)
    assert np.abs(rhs - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("term", ["kerr", "photo", "defoc", "ste"])
def test_split_term_on_a_general_field_is_its_tone_multiplier(term):
    # On a tone (constant |u|, constant densities) a term's spectral operator shows as one
    # complex factor M[m] per frequency bin over its pointwise source: |u|^2 u (Kerr),
    # rho u (plasma defocusing), rho_s u (STE index) and, for photo-ionisation, photo u,
    # whose alpha u = photo u part is added unfiltered. On a general field, with random
    # densities, each term must be exactly the diagonal operator with those factors applied
    # to its source. An operator applied before the pointwise product agrees on tones and
    # fails here.
    on = {"kerr": dict(enable_kerr_instantaneous=True),
          "photo": dict(enable_photoionization_loss=True),
          "defoc": dict(enable_plasma_defocusing=True),
          "ste": dict(enable_ste=True, enable_ste_index=True)}[term]
    cfg = _small_cfg(**{**_NO_NL, **on, "enable_self_steepening": True,
                        "enable_space_time_focusing": True, "enable_spectral_filter": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)

    def source(u, a, rho, rho_s):
        return {"kerr": np.abs(u) ** 2 * u, "photo": a * u, "defoc": rho * u,
                "ste": rho_s * u}[term]

    flat = np.full(g["rr"].shape, 1e20)
    M = np.zeros(cfg.Nt, complex)
    for m in range(cfg.Nt):
        u = _tone(g, m, 4e13)
        rhs, a = nl.split(u, flat, flat)
        sv = source(u, a, flat, flat)
        ratio = (sv - rhs) / sv if term == "photo" else rhs / sv
        M[m] = ratio[0, 0]
        assert np.allclose(ratio, M[m], rtol=1e-12, atol=1e-12), m     # one factor per tone
    assert np.ptp(np.abs(M)) > 1e-3 * np.abs(M).max()   # the operators are not a constant
    u = _field(cfg, g, seed=5, scale=20.0)
    rng = np.random.default_rng(9)
    rho, rho_s = rng.uniform(0, 1e20, u.shape), rng.uniform(0, 1e20, u.shape)
    rhs, a = nl.split(u, rho, rho_s)
    sv = source(u, a, rho, rho_s)
    assert np.abs(sv).max() > 0
    pred = np.fft.ifft(np.fft.fft(sv, axis=1) * M, axis=1)
    if term == "photo":
        pred = sv - pred
    lit = np.abs(u) ** 2 >= 1e-30
    assert np.abs(rhs[lit] - pred[lit]).max() <= 1e-11 * np.abs(pred[lit]).max()


def test_split_raman_of_constant_intensity():
    # Raman alone on a constant |u|^2 = A: the response integral is the DC value of R_f
    # (its closed form is pinned above), rhs = i kappa f_R A Re(R_f[0]) u
    cfg = _small_cfg(**{**_NO_NL, "enable_kerr_raman": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _tone(g, 0, 3e13)
    rhs, _ = nl.split(u, np.zeros(u.shape), None)
    A = np.abs(u[0, 0]) ** 2
    ref = 1j * _kappa(cfg) * cfg.f_R * A * g["R_f"][0].real * u
    assert 0.9 < g["R_f"][0].real < 1.1          # unit-area response, sampled
    assert np.abs(rhs - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("m", [3, 11])
def test_split_raman_of_a_modulated_intensity(m):
    # Raman alone (no spectral operators) on |u|^2 = a2 (1 + cos(theta n) / 2), u real:
    # the delayed response Re ifft(fft(|u|^2) R_f) of one cosine is
    # a2 Re R_f[0] + (a2 / 4) Re(R_f[m] e^(i theta n) + R_f[-m] e^(-i theta n)), so
    # rhs = i kappa f_R (that) u. Unlike the constant-intensity case this sees R_f away from
    # DC: a time-reversed (conjugated) response or the imaginary part fails it.
    cfg = _small_cfg(**{**_NO_NL, "enable_kerr_raman": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    Nt = cfg.Nt
    th = 2 * np.pi * m * np.arange(Nt) / Nt
    a2 = 3e13 / g["invE2"]
    u = np.broadcast_to(np.sqrt(a2 * (1 + 0.5 * np.cos(th))), g["rr"].shape).astype(complex)
    rhs, a = nl.split(u, np.zeros(u.shape), None)
    R = g["R_f"]
    resp = a2 * R[0].real + (a2 / 4) * (R[m] * np.exp(1j * th) + R[-m] * np.exp(-1j * th)).real
    ref = 1j * _kappa(cfg) * cfg.f_R * resp * u
    assert np.all(a == 0) and abs(R[m] - np.conj(R[-m])) <= 1e-12 * abs(R[m])  # real response
    assert abs(np.angle(R[m])) > 1e-2                 # the response is delayed at f_m
    assert np.abs(rhs - ref).max() <= 1e-12 * np.abs(ref).max()


def test_split_raman_of_a_general_intensity_is_the_time_domain_convolution():
    # Raman alone (no spectral operators) on a general field: the response is the circular
    # convolution of |u|^2 with the sampled causal kernel h[k] = Omega_r^2 tau_s dt
    # exp(-k dt / tau_d) sin(k dt / tau_s), k = 1 .. N/2 - 1 (the lags ifftshift puts at
    # t > 0), summed here directly in the time domain, with no FFT: rhs = i kappa f_R
    # (h * |u|^2) u. Pins the whole Raman path on arbitrary intensities (orientation,
    # lag range, the dt factor and the real part), not only per Fourier component.
    cfg = _small_cfg(N=9, **{**_NO_NL, "enable_kerr_raman": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    Nr, Nt = g["rr"].shape
    rng = np.random.default_rng(11)
    a2 = 3e13 / g["invE2"]
    amp = np.sqrt(a2 * rng.uniform(0.0, 1.0, (Nr, Nt)))
    u = amp * np.exp(2j * np.pi * rng.uniform(size=(Nr, Nt)))
    rhs, a = nl.split(u, np.zeros(u.shape), None)
    dt = g["dt"]
    k = np.arange(1, Nt // 2)
    h = cfg.Omega_r2 * cfg.tau_s * dt * np.exp(-k * dt / cfg.tau_d) * np.sin(k * dt / cfg.tau_s)
    I = np.abs(u) ** 2
    resp = np.zeros_like(I)
    for j, kk in enumerate(k):
        resp += h[j] * np.roll(I, kk, axis=1)          # I[n - kk]
    ref = 1j * _kappa(cfg) * cfg.f_R * resp * u
    assert np.all(a == 0)
    assert np.abs(rhs - ref).max() <= 1e-11 * np.abs(ref).max()
    # a time-reversed kernel (I[n + k]) is a different operator on this field
    rev = np.zeros_like(I)
    for j, kk in enumerate(k):
        rev += h[j] * np.roll(I, -kk, axis=1)
    assert np.abs(rev - resp).max() > 1e-2 * np.abs(resp).max()


def test_split_ste_index_term_alone():
    # the STE index change alone: extra = +i ste_pref rho_s u, so rhs = i ste_pref rho_s u
    # (ste_pref itself is pinned by the recorded delta n_STE above)
    cfg = _small_cfg(**{**_NO_NL, "enable_ste": True, "enable_ste_index": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g, scale=10.0)
    rho_s = np.random.default_rng(6).uniform(0, 1e20, u.shape)
    rhs, a = nl.split(u, np.zeros(u.shape), rho_s)
    assert g["ste_pref"] > 0 and np.all(a == 0)
    ref = 1j * g["ste_pref"] * rho_s * u
    assert np.abs(rhs - ref).max() <= 1e-12 * np.abs(ref).max()


def test_split_zero_field_samples_have_zero_rhs():
    # rhs is set to 0 where |u|^2 < 1e-30 (operators.py:145): with self-steepening the
    # spectral terms spread into the field's gaps, so without the mask rhs would not vanish there
    cfg = _small_cfg(**{**_NO_NL, "enable_kerr_instantaneous": True,
                        "enable_self_steepening": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g, scale=10.0)
    u[:, : cfg.Nt // 4] = 0.0
    rhs, _ = nl.split(u, np.zeros(u.shape), None)
    assert np.all(rhs[:, : cfg.Nt // 4] == 0)
    lit = np.abs(u[:, cfg.Nt // 2]) ** 2 > 1e-20          # rows the beam reaches
    assert lit.sum() > 10 and np.abs(rhs[lit, cfg.Nt // 2]).min() > 0


def test_split_alpha_is_clipped_at_zero():
    # alpha = max(nan_to_num(photo + plasma_pref rho), 0) (operators.py:98-107): a negative
    # density (outside the physical range) gives alpha = 0, not a gain
    cfg = _small_cfg(**{**_NO_NL, "enable_plasma_absorption": True})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g)
    rho = np.full(u.shape, -1e20)
    rho[0] = 1e20
    _, a = nl.split(u, rho, None)
    assert np.all(a[1:] == 0)
    assert np.allclose(a[0], g["sigmaomega"] / 2 * 100 * 1e20, rtol=1e-14, atol=0)


def test_split_terms_add_across_toggles():
    # split() sums independent terms (operators.py:89-146): the Kerr and Raman parts of kerr_I,
    # the photo term, the plasma term (coefficient [absorb] + [defoc] (phase - 1)), the STE
    # index term, and alpha u with alpha = photo + [absorb] plasma_pref rho (>= 0, so the
    # clip is inert). With self-steepening, space-time focusing and the spectral filter on,
    # T and 1/U mix the frequencies of a general field, and still the rhs with every term on
    # is the sum of the rhs with one term on. This pins how the terms combine, which the
    # one-term pins above leave open (a term that overwrites instead of adding fails).
    ops = dict(enable_self_steepening=True, enable_space_time_focusing=True,
               enable_spectral_filter=True)
    terms = [dict(enable_kerr_instantaneous=True), dict(enable_kerr_raman=True),
             dict(enable_photoionization_loss=True), dict(enable_plasma_absorption=True),
             dict(enable_plasma_defocusing=True), dict(enable_ste=True, enable_ste_index=True)]
    every = {k: v for t in terms for k, v in t.items()}
    cfg = _small_cfg(**{**_NO_NL, **ops, **every})
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g, seed=4, scale=40.0)           # peak 2.5e14 W/cm^2: photo term alive
    rng = np.random.default_rng(8)
    rho = rng.uniform(0, 1e20, u.shape)
    rho_s = rng.uniform(0, 1e20, u.shape)
    rhs_all, a_all = nl.split(u, rho, rho_s)
    parts, alphas = [], []
    for t in terms:
        c1 = _small_cfg(**{**_NO_NL, **ops, **t})
        g1 = O.build_grids(c1)
        for key in ("T_op", "inv_U_nl", "R_f", "sigmaomega", "komega"):
            assert np.array_equal(g1[key], g[key]), key
        r1, a1 = O.NonlinearOperator(c1, g1).split(u, rho, rho_s)
        parts.append(r1)
        alphas.append(a1)
    scale = sum(np.abs(r).max() for r in parts)
    for t, r in zip(terms, parts):                   # every term matters at this tolerance
        assert np.abs(r).max() > 1e-6 * scale, t
    assert np.abs(rhs_all - sum(parts)).max() <= 1e-12 * scale
    assert np.abs(a_all - sum(alphas)).max() <= 1e-12 * np.abs(a_all).max()
    assert np.abs(a_all).max() > 0 and np.abs(alphas[2]).max() > 0 and np.abs(alphas[3]).max() > 0


def _pchip_slopes(x, y):
    """Knot derivatives of the monotone piecewise-cubic Hermite interpolant as published
    (Fritsch & Carlson 1980, in the form of Moler, Numerical Computing with MATLAB, 3.4):
    0 where the secants change sign or vanish, else the weighted harmonic mean of the two
    secants with weights 2 h_k + h_(k-1) and h_k + 2 h_(k-1); at each end the one-sided
    three-point formula, set to 0 if it opposes the first secant and limited to 3 times
    that secant where the first two secants differ in sign."""
    h = np.diff(x)
    m = np.diff(y) / h
    d = np.zeros_like(y)
    for k in range(1, len(x) - 1):
        if m[k - 1] * m[k] > 0:
            w1, w2 = 2 * h[k] + h[k - 1], h[k] + 2 * h[k - 1]
            d[k] = (w1 + w2) / (w1 / m[k - 1] + w2 / m[k])

    def end(h0, h1, m0, m1):
        e = ((2 * h0 + h1) * m0 - h0 * m1) / (h0 + h1)
        if np.sign(e) != np.sign(m0):
            return 0.0
        if np.sign(m0) != np.sign(m1) and abs(e) > abs(3 * m0):
            return 3 * m0
        return e
    d[0] = end(h[0], h[1], m[0], m[1])
    d[-1] = end(h[-1], h[-2], m[-1], m[-2])
    return d


def test_pchip_is_the_published_monotone_interpolant():
    # the photo-ionisation rate W(I) is a PCHIP through the Keldysh knots (keldysh.py:144-150);
    # the oracle and the host tables use SciPy's PchipInterpolator where the reference uses
    # cupyx's. Both follow the published Fritsch-Carlson slopes: pin SciPy's knot derivatives
    # to that rule on the real knots (both channels) and on a small set with a sign change,
    # a flat run and an overshooting end.
    g = O.build_grids(_small_cfg())
    k = g["keldysh"]
    for W in (k["W_knots_g"], k["W_knots_s"]):
        x = k["f_spline"].x
        sp = O.PchipInterpolator(x, W)
        ref = _pchip_slopes(x, W)
        got = np.append(sp.c[2], sp.derivative()(x[-1]))
        assert np.allclose(got, ref, rtol=1e-10, atol=1e-300)
        assert np.allclose(sp(x), W, rtol=1e-12, atol=0)
    x = np.array([0.0, 1.0, 1.5, 3.0, 4.0, 4.2, 6.0])
    y = np.array([0.0, 2.0, 2.0, 1.0, 3.5, 3.6, 3.0])
    sp = O.PchipInterpolator(x, y)
    ref = _pchip_slopes(x, y)
    assert ref[1] == 0 and ref[2] == 0 and ref[3] == 0   # flat run, then a sign change
    got = np.append(sp.c[2], sp.derivative()(x[-1]))
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-15)


def _bessel_series(n, x, digits=80):
    """J_n(x) = sum_m (-1)^m (x/2)^(2m+n) / (m! (m+n)!), summed in `digits`-digit decimal
    arithmetic (no cancellation loss for the arguments used here)."""
    from decimal import Decimal, localcontext
    with localcontext() as ctx:
        ctx.prec = digits
        h = Decimal(repr(float(x))) / 2
        term = h ** n
        for k in range(1, n + 1):
            term /= k
        total, m = term, 0
        while abs(term) > Decimal(10) ** (-digits + 10) or m < 5:
            m += 1
            term = -term * h * h / (m * (m + n))
            total += term
        return float(total)


def test_bessel_functions_are_the_power_series():
    # the Hankel matrix Y (grids.py) is built from j0 / j1 and the zeros of J0; the oracle
    # and the host use SciPy's, the reference cupyx's. Pin SciPy's values to the defining
    # series: the first 20 zeros of J0 (arguments up to 62), and j0 / j1 at spread points.
    z = O.jn_zeros(0, 20)
    for zk in z:
        assert abs(_bessel_series(0, zk)) < 1e-14
        assert abs(O.j1(zk) - _bessel_series(1, zk)) < 1e-14
    assert np.all(np.diff(z) > 3.0) and abs(z[0] - 2.404825557695773) < 1e-15
    for x in np.linspace(0.05, 60.0, 37):
        assert abs(O.j0(x) - _bessel_series(0, x)) < 1e-14, x
        assert abs(O.j1(x) - _bessel_series(1, x)) < 1e-14, x


@pytest.mark.parametrize("N", [64, 512, 4096])
def test_numpy_fft_is_the_dft(N):
    # the oracle's transforms along t are numpy.fft where the reference calls cupy.fft:
    # pin numpy's forward and inverse transforms (sign, ordering, 1/N on the inverse) to the
    # direct DFT summed in long double, with exp(-2 pi i m / N) at m = k n mod N
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    n = np.arange(N)
    ang = (2 * np.pi / np.longdouble(N)) * n.astype(np.longdouble)
    w = np.cos(ang) - 1j * np.sin(ang).astype(np.longdouble)
    ks = [0, 1, 2, N // 3, N // 2, N - 1] if N > 64 else list(range(N))
    xr, xi = x.real.astype(np.longdouble), x.imag.astype(np.longdouble)
    for k in ks:
        m = (k * n) % N
        c, s_ = w.real[m], w.imag[m]
        ref = complex(float(np.sum(xr * c - xi * s_)), float(np.sum(xr * s_ + xi * c)))
        assert abs(np.fft.fft(x)[k] - ref) <= 1e-14 * np.sqrt(N) * np.abs(x).max(), k
    X = np.fft.fft(x)
    assert np.abs(np.fft.ifft(X) - x).max() <= 1e-14 * np.abs(x).max()
    assert np.allclose(np.fft.ifft(X * N)[:3], [np.sum(X * np.exp(2j * np.pi * np.arange(N) * t / N))
                                                 for t in range(3)], rtol=1e-12, atol=0)


def _bessel_asymptotic(nu, x):
    """Hankel's expansion (Abramowitz & Stegun 9.2.5-9.2.10): J_nu(x) = sqrt(2 / (pi x))
    (P cos chi - Q sin chi), chi = x - (nu / 2 + 1 / 4) pi, with P, Q the even and odd terms
    of sum_k (-1)^floor(k/2) a_k(nu) / x^k, a_k = prod_{j<=k} (4 nu^2 - (2j - 1)^2) / (k! 8^k),
    summed until the terms fall below 1e-20 (x >= 62). cos chi and sin chi come from cos x
    and sin x of the float argument, whose range reduction is exact."""
    mu = 4.0 * nu * nu
    a, P, Q, k = 1.0, 1.0, 0.0, 0
    while abs(a) >= 1e-20 or k < 3:
        k += 1
        a *= (mu - (2 * k - 1) ** 2) / (k * 8.0 * x)
        if k % 2 == 0:
            P += (-1) ** (k // 2) * a
        else:
            Q += (-1) ** (k // 2) * a
    c, s_ = np.cos(x), np.sin(x)
    r = np.sqrt(0.5)
    cc, ss = ((c + s_) * r, (s_ - c) * r) if nu == 0 else ((s_ - c) * r, -(s_ + c) * r)
    return np.sqrt(2.0 / (np.pi * x)) * (P * cc - Q * ss)


def test_bessel_functions_at_large_arguments():
    # the Hankel matrix at N = 1025 evaluates j0 / j1 up to j_N = 3219.3: pin SciPy's values
    # there to Hankel's asymptotic expansion (the power series pin above stops at 62)
    z = O.jn_zeros(0, 1025)
    # McMahon's expansion of the zeros (A&S 9.5.12, nu = 0): b = (s - 1/4) pi,
    # j_s = b + 1/(8b) - 124/(3 (8b)^3) + 120928/(15 (8b)^5) - ...
    b = (np.arange(1, 1026) - 0.25) * np.pi
    mc = b + 1 / (8 * b) - 124 / (3 * (8 * b) ** 3) + 120928 / (15 * (8 * b) ** 5)
    assert np.abs(z[19:] - mc[19:]).max() < 1e-11
    for x in np.concatenate([z[19::16], np.linspace(62.0, 3300.0, 97)]):
        assert abs(O.j0(x) - _bessel_asymptotic(0, x)) < 1e-14, x
        assert abs(O.j1(x) - _bessel_asymptotic(1, x)) < 1e-14, x


def test_alpha_pointwise_closed_form():
    cfg = _small_cfg()
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    u = _field(cfg, g, scale=30.0)
    rho = np.random.default_rng(3).uniform(0, 1e21, u.shape)
    _, a = nl.split(u, rho, None)
    I = np.abs(u) ** 2 * g["invE2"]
    W = g["keldysh"]["f_spline"](np.clip(I, 1.0, 9.99e16))
    photo = W * 1e6 * cfg.Ui / (cfg.n0 * c * epsilon_0 * (np.abs(u) ** 2 + 1e-30)) * (1 - rho / cfg.rho_max)
    ref = photo + g["sigmaomega"] / 2 * 100 * rho
    assert np.abs(a - ref).max() <= 1e-12 * ref.max()


# ---------------------------------------------------------------- rate equation
def _scalar_rate(I, W, Ws, dt, bg, bs, na, itr, its, ste):
    """Independent scalar restatement of kernels.py:48-105 on a given (I, W, Ws) row."""
    import math
    ne = ns = 0.0
    NE, NS = [0.0], [0.0]

    def step(x, S, L):
        Ldt = L * dt
        p = (math.exp(Ldt) - 1.0) / Ldt if abs(Ldt) > 1e-6 else 1.0 + Ldt * (0.5 + Ldt * (1 / 6 + Ldt / 24))
        xn = math.exp(Ldt) * x + S * dt * p
        return xn if (math.isfinite(xn) and xn >= 0) else 0.0
    for k in range(len(I) - 1):
        Ia, Wa, Wsa = 0.5 * (I[k] + I[k + 1]), 0.5 * (W[k] + W[k + 1]), 0.5 * (Ws[k] + Ws[k + 1])
        depl = min(max(1.0 - ne / na, 0.0), 1.0)
        Se = Wa * depl + ((Wsa + bs * Ia * ne) * (ns / na) if ste else 0.0)
        Le = bg * Ia * depl - itr
        Ss = itr * ne if ste else 0.0
        Ls = (-(Wsa + bs * Ia * ne) / na - its) if ste else 0.0
        a, b = step(ne, Se, Le), (step(ns, Ss, Ls) if ste else 0.0)
        if a + b > na:
            s = na / (a + b)
            a, b = a * s, b * s
        ne, ns = a, b
        NE.append(ne)
        NS.append(ns)
    return np.array(NE), np.array(NS)


@pytest.mark.parametrize("ste", [False, True])
def test_rate_rows_vs_scalar_restatement(ste):
    cfg = _small_cfg(Nt=96, N=5, enable_ste=ste, tau_ste=1e-12 if ste else None)
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    rng = np.random.default_rng(7)
    tt = g["tlist"]
    amp = np.array([0.5, 1.0, 2.0, 4.0])[:, None] * 6e13
    I = amp * np.exp(-2 * tt ** 2 / (60e-15) ** 2)[None, :] * (1 + 0.05 * rng.standard_normal((4, cfg.Nt)))
    u = np.sqrt(np.abs(I) / g["invE2"]).astype(complex)
    rho, rho_s = np.zeros(u.shape), np.zeros(u.shape)
    nl.update_plasma(u, rho, rho_s, g["dt"])
    kel = g["keldysh"]
    lut = (kel["NLUT"], kel["logImin"], kel["inv_dlog"], kel["Imin"], kel["Imax"])
    for r in range(4):
        Ir = (u[r].real ** 2 + u[r].imag ** 2) * g["invE2"]
        W = O.lut_rate(Ir, kel["W_LUT"], *lut)
        Ws = O.lut_rate(Ir, kel["W_LUT_s"], *lut) if ste else np.zeros_like(Ir)
        NE, NS = _scalar_rate(Ir, W, Ws, g["dt"], nl.avalanche, nl.avalanche_s, nl.rho_max,
                              nl.inv_taur, nl.inv_tau_ste, ste)
        assert np.allclose(rho[r], NE, rtol=1e-13, atol=0)
        assert np.allclose(rho_s[r], NS, rtol=1e-13, atol=0)
    assert rho.max() > 1e15                       # the case exercises the rates
    assert np.all(rho >= 0) and np.all(rho + rho_s <= cfg.rho_max * (1 + 1e-15))


def test_rate_constant_intensity_exact_solution():
    # Frozen coefficients: no depletion (huge rho_max) and constant I make the
    # exact-exponential step exact: n(t) = (W/L)(exp(L t) - 1), L = beta I - 1/tau_r.
    cfg = _small_cfg(Nt=64, N=3, rho_max=1e300)
    g = O.build_grids(cfg)
    nl = O.NonlinearOperator(cfg, g)
    I0 = 3e13
    u = np.full((2, cfg.Nt), np.sqrt(I0 / g["invE2"]), complex)
    rho, rho_s = np.zeros(u.shape), np.zeros(u.shape)
    nl.update_plasma(u, rho, rho_s, g["dt"])
    kel = g["keldysh"]
    W = O.lut_rate(np.array([I0]), kel["W_LUT"], kel["NLUT"], kel["logImin"], kel["inv_dlog"],
                   kel["Imin"], kel["Imax"])[0]
    Iq = (u[0, 0].real ** 2 + u[0, 0].imag ** 2) * g["invE2"]
    L = nl.avalanche * Iq - nl.inv_taur
    t = np.arange(cfg.Nt) * g["dt"]
    ref = W / L * np.expm1(L * t)
    assert np.allclose(rho[0], ref, rtol=1e-11, atol=0)


# ---------------------------------------------------------------- whole loop
def test_linear_only_run_follows_gaussian_beam_and_z_labels():
    # All six field toggles off, STF and spectral filter off: the loop is linear
    # and fluence obeys the Gaussian-beam law; the state recorded with label
    # z_i = begin + i dz has propagated (i+1) dz (integrator.py:153-159).
    cfg = _small_cfg(nz=3, begin=-4e-3, end=-1e-3, N=257, Nt=64, R_factor=15,
                     enable_kerr_instantaneous=False, enable_kerr_raman=False,
                     enable_self_steepening=False, enable_photoionization_loss=False,
                     enable_plasma_defocusing=False, enable_plasma_absorption=False,
                     enable_space_time_focusing=False, enable_spectral_filter=False,
                     rho_t_stride=0)
    it = O.Integrator(cfg)
    res = it.propagate()
    assert len(res["z"]) == cfg.nz + 1
    assert np.allclose(res["z"], cfg.begin + np.arange(cfg.nz + 1) * cfg.dz)
    g, w0 = it.g, cfg.w0
    F0 = np.sum(np.abs(np.exp(-g["tlist"] ** 2 / (cfg.delta_t / np.sqrt(2 * np.log(2))) ** 2)) ** 2) \
        * g["dt"] * g["invE2"] * cfg.E0 ** 2
    for i, z in enumerate(res["z"]):
        zp = cfg.begin + (i + 1) * cfg.dz
        q = 1 + 2j * zp / cfg.b
        ref = F0 / abs(q) ** 2 * np.exp(-2 * g["rlist"] ** 2 * (1 / (w0 ** 2 * q)).real)
        Nr = g["rlist"].size
        got = res["fluence_rz"][i, Nr:]
        assert np.abs(got - ref).max() < 2e-6 * ref.max()


def test_step_ends_with_the_radial_absorber():
    # Integrator.step multiplies the stepped field by mask_r = exp(-(r / 0.9 R)^20)
    # (integrator.py:148, grids.py:138). A flat field reaching the grid edge and a 1 nm step
    # (linear-only: the propagation is the identity to ~1e-5) must come back as u0 mask_r.
    cfg = _small_cfg(nz=1, begin=0.0, end=1e-9, **{**_NO_NL})
    flat = lambda rr, tt, cfg_, g_: np.full(rr.shape, 1e8 + 0j)   # noqa: E731
    it = O.Integrator(cfg, envelope=flat)
    u0 = np.array(it.u, copy=True)
    it.step(it.dz)
    r = it.g["rlist"]
    R = it.g["R"]
    mask = np.exp(-(r / (0.9 * R)) ** 20)
    assert mask[-1] < 0.2 and mask[0] == 1.0        # the absorber matters at the edge
    ref = u0 * mask[:, None]
    assert np.abs(it.u - ref).max() <= 1e-4 * np.abs(ref).max()


def test_step_alpha_half_steps_total_one_absorption_length():
    # step() applies exp(-alpha dz / 2) before and after the RK4 stage (integrator.py:141-147).
    # With plasma absorption alone on a uniform density (T = U^-1 = 1, so the RK4 rhs is 0 and
    # alpha commutes with the linear steps) a step is the linear-only step times exp(-alpha dz)
    kw = dict(nz=1, begin=0.0, end=5e-6)
    it_a = O.Integrator(_small_cfg(**kw, **{**_NO_NL, "enable_plasma_absorption": True}))
    it_l = O.Integrator(_small_cfg(**kw, **_NO_NL))
    for it in (it_a, it_l):
        it.rho[:] = 1e20
        it.step(it.dz)
    alpha = it_a.g["sigmaomega"] / 2 * 100 * 1e20
    assert alpha * it_a.dz > 1e-3                   # a visible attenuation
    ref = it_l.u * np.exp(-alpha * it_a.dz)
    assert np.abs(it_a.u - ref).max() <= 1e-12 * np.abs(ref).max()


def test_results_mirror_the_radial_axis():
    # _results mirrors every (z, r) map as hstack([a[:, ::-1], a]) and r as [-r[::-1], r]
    # (integrator.py:238-298): symmetric about the centre, the axis odd
    cfg = _small_cfg(nz=2, end=20e-6, N=33, Nt=64, rho_t_stride=0)
    res = O.Integrator(cfg).propagate()
    Nr = cfg.N - 1
    assert np.array_equal(res["r"], np.concatenate([-res["rlist"][::-1], res["rlist"]]))
    for k in ("fluence_rz", "rho_rz", "Ipeak_rz", "absorbed_rz"):
        a = res[k]
        assert a.shape[1] == 2 * Nr, k
        assert np.array_equal(a[:, :Nr], a[:, Nr:][:, ::-1]), k
        assert np.abs(a[:, Nr]).max() >= np.abs(a[:, -1]).max(), k   # centre outward


def test_lut_rate_index_and_interpolation():
    # interpolate_pi_rate (kernels.py:14-26): fidx = (log10 I - log_min) inv_step, truncated;
    # at the table's own sample points it returns the table, and at i + 3/4 it interpolates
    # from sample i (rounding would interpolate from i + 1, off by the table's curvature)
    cfg = _small_cfg()
    g = O.build_grids(cfg)
    k = g["keldysh"]
    tab, n = k["W_LUT"], k["NLUT"]
    lmin, inv = k["logImin"], k["inv_dlog"]
    i = np.array([3, 100, 2047, n - 3])
    I_at = 10.0 ** (lmin + (i + 1e-9) / inv)          # just above sample i
    I_mid = 10.0 ** (lmin + (i + 0.75) / inv)
    args = (n, lmin, inv, k["Imin"], k["Imax"])
    assert np.allclose(O.lut_rate(I_at, tab, *args), tab[i], rtol=1e-6, atol=0)
    assert np.allclose(O.lut_rate(I_mid, tab, *args), tab[i] + 0.75 * (tab[i + 1] - tab[i]),
                       rtol=1e-9, atol=0)
    # the table spans the PCHIP knots' decades uniformly in log10 I (4096 points, 4095 steps)
    assert n == 4096 and abs(inv * (np.log10(9.99e16) - np.log10(1.0)) - 4095) < 1e-9


def test_record_invariants():
    cfg = _small_cfg(nz=2, end=20e-6, N=65, Nt=128, energy_uJ=40.0, w0=40e-6,
                     absorb_time_bins_fs=(-400, -50, 0, 50, 400),
                     rho_snapshot_t_fs=20.0, rho_t_stride=3, rho_r_stride=2)
    res = O.Integrator(cfg).propagate()
    tot = res["absorbed_rz"]
    bins = res["absorbed_rz_bins"].sum(axis=1)
    assert np.abs(bins - tot).max() <= 1e-6 * np.abs(tot).max()
    assert np.all(res["rho_rz_at"] <= res["rho_rz"])
    assert np.isclose(res["Ipeak_rz"].max(), res["Imax_z"].max(), rtol=1e-6)
    assert res["rho_rz"].max() > 0
    assert res["rho_rzt"].shape == (3, 32, 43)


def test_rate_recurrence_reproduces_the_reference_0d_vectors():
    # The carrier kernel's recurrence (kernels.py:59-104) and the reference's own 0-D
    # re-integration (figures_article.py:161-194) are the same two-population scheme: with
    # inv_tau_ste = 0 and the rates evaluated at the samples they differ only in
    # expm1(x)/x against (exp(x)-1)/x. The reference-produced vectors (tests/golden/
    # rate0d_ste.npz: STE on, all three avalanche / trapping variants, reaching the
    # n_e + n_s = n_a rescaling) therefore pin the STE branch, the depletion factor and
    # the rescaling of the oracle's recurrence.
    import os
    from oracle import postprocess_numpy as P0
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "rate0d_ste.npz"))
    prm = P0.Params0D()
    I = np.clip(gold["I"], 0.0, None)
    W, Ws = prm.W_vb(I), prm.W_ste(I)
    dt = (gold["t_fs"][1] - gold["t_fs"][0]) * 1e-15
    assert np.allclose(np.diff(gold["t_fs"]) * 1e-15, dt, rtol=1e-12, atol=0)
    for name, bg, itr in (("full", prm.beta_g, 1.0 / prm.tau_r),
                          ("noav", 0.0, 1.0 / prm.tau_r),
                          ("notrap", prm.beta_g, 0.0)):
        ne, ns = np.zeros_like(I), np.zeros_like(I)
        O.rate_recurrence_rows(I, W, Ws, ne, ns, dt, bg, prm.beta_s, prm.rho_max, itr, 0.0, True)
        scale = max(np.abs(gold["ne_" + name]).max(), np.abs(gold["ns_" + name]).max())
        assert np.abs(ne - gold["ne_" + name]).max() <= 1e-9 * scale, name
        assert np.abs(ns - gold["ns_" + name]).max() <= 1e-9 * scale, name
    assert np.max(gold["ne_full"] + gold["ns_full"]) >= prm.rho_max * (1 - 1e-12)


def test_rate_equation_rows_depletion_closed_form():
    # With no avalanche and no trapping (L = 0) the exact-exponential step is n + S dt, and
    # the depletion factor gives n_{k+1} = n_k + W dt (1 - n_k / n_a): the closed form
    # n_a (1 - (1 - W dt / n_a)^k), through rate_equation_rows itself (field -> intensity
    # -> rate lookup -> recurrence). Fails with the depletion factor dropped or taken from
    # the new density, or with the average of the wrong samples.
    cfg = _small_cfg(Nt=512, N=3)
    g = O.build_grids(cfg)
    kel = g["keldysh"]
    lut = (kel["NLUT"], kel["logImin"], kel["inv_dlog"], kel["Imin"], kel["Imax"])
    I0 = 8e13
    E = np.full((2, cfg.Nt), np.sqrt(I0 / g["invE2"]), complex)
    W = O.lut_rate(np.array([I0]), kel["W_LUT"], *lut)[0]
    dt, na = g["dt"], 2.1e22
    a = W * dt / na
    k = np.arange(cfg.Nt)
    assert a * cfg.Nt > 1.0                                       # depletion really acts
    ne, ns = np.zeros(E.shape), np.zeros(E.shape)
    O.rate_equation_rows(E, ne, ns, dt, g["invE2"], 0.0, 0.0, na, 0.0, 0.0, False,
                         kel["W_LUT"], kel["W_LUT_s"], *lut)
    exact = na * -np.expm1(k * np.log1p(-a))
    assert np.abs(ne[0] - exact).max() <= 1e-11 * na
    assert not ns.any()


def _balance_run(dz=20e-6, n=31, rho0=1e17, w0=100e-6, R_factor=10, energy_uJ=2.0, **toggles):
    """n steps at a frozen uniform density, each followed by a record (integrator.py:
    138-149, 179-224); returns the integrator and its results."""
    cfg = _small_cfg(nz=n - 1, begin=0.0, end=(n - 1) * dz, N=129, Nt=128, R_factor=R_factor,
                     w0=w0, energy_uJ=energy_uJ, rho_t_stride=0, save_stride=1,
                     **{**_NO_NL, **toggles})
    it = O.Integrator(cfg)
    for i in range(n):
        it.rho[:] = rho0
        it.step(it.dz)
        it._record(cfg.begin + i * it.dz)
    return it, it._results()


def _energies_uJ(it, fluence_rz):
    r = it.g["rlist"]
    w = 2.0 * np.pi * r * np.diff(r, prepend=0.0)
    return (fluence_rz[:, r.size:].astype(np.float64) * w).sum(axis=1) * 1e10


def _trapz_over_U0(it, d):
    z = it.z_saved[:it.k_save]
    return np.concatenate([[0.0], np.cumsum(0.5 * (d[:-1] + d[1:]) * np.diff(z))]) / it.U0_uJ


def test_energy_bookkeeping_balances_the_field_energy():
    # integrator.py:199-212, 226-236. With T = U^-1 = 1 the loss terms of the RK4 rhs
    # cancel alpha u, so the field loses energy only through the alpha half-steps. The
    # bookkeeping must then account for what the field lost: the cumulative E_*_z
    # (trapezoid of dE/dz over z, over U0) equals (U(z_0) - U(z_k)) / U0 with U from the
    # recorded fluence (sum_r F 2 pi r dr, J/cm^2 x cm^2 -> uJ). Plasma absorption on a
    # uniform density (alpha uniform): dE/dz = 2 alpha U and absorbed = 2 alpha 100 F per
    # r, and the balance holds to the trapezoid error; photo-ionisation (alpha peaked in
    # r and t): to the O(dz) error of evaluating the loss at the records.
    it, res = _balance_run(enable_plasma_absorption=True)
    alpha = it.g["sigmaomega"] / 2 * 100 * 1e17
    U = _energies_uJ(it, res["fluence_rz"])
    n = it.k_save
    d = it._dEdz_plasma_uJm[:n]
    bal = (U[0] - U) / it.U0_uJ
    E = res["E_plasma_z"]
    assert bal[-1] > 0.02                                    # a visible loss
    assert np.abs(d - 2 * alpha * U).max() <= 1e-6 * d.max()
    F = res["fluence_rz"][:, it.g["rlist"].size:].astype(np.float64)
    ab = res["absorbed_rz"][:, it.g["rlist"].size:]
    assert np.abs(ab - 2 * alpha * 100 * F).max() <= 1e-6 * ab.max()
    assert np.abs(E - bal).max() <= 1e-4 * bal[-1]
    assert np.array_equal(res["E_total_z"], res["E_MPI_z"] + res["E_plasma_z"])
    assert not res["E_MPI_z"].any()
    # the same checks reject each plausible slip in the bookkeeping
    r = it.g["rlist"]
    for bad in (2.0 * d,                                               # x2 on dE/dz
                d / (2.0 * np.pi * r).mean(),                          # no 2 pi r weight
                d * 100.0):                                            # x100 on dE/dz
        assert np.abs(_trapz_over_U0(it, bad) - bal).max() > 0.1 * bal[-1]
    assert np.abs(ab / 100.0 - 2 * alpha * 100 * F).max() > 0.5 * ab.max() / 100.0

    it, res = _balance_run(enable_photoionization_loss=True, w0=20e-6, R_factor=4, energy_uJ=10.0)
    U = _energies_uJ(it, res["fluence_rz"])
    bal = (U[0] - U) / it.U0_uJ
    E = res["E_MPI_z"]
    assert bal[-1] > 5e-3 and not res["E_plasma_z"].any()
    assert abs(E[-1] - bal[-1]) <= 1e-2 * bal[-1]
    # absorbed_rz and dE/dz are the same per-row loss integral (x100, and x 2 pi r dr 1e10)
    r = it.g["rlist"]
    ab = res["absorbed_rz"][:, r.size:].astype(np.float64)
    w = 2.0 * np.pi * r * np.diff(r, prepend=0.0)
    d = it._dEdz_photo_uJm[:it.k_save]
    assert np.abs((ab / 100.0 * w).sum(axis=1) * 1e10 - d).max() <= 1e-6 * d.max()


def test_keldysh_monotone_envelope():
    # keldysh.py:132-148: rate(monotone=True) is the PCHIP (in log10 I) of the running
    # maximum of rate_raw on 6000 points over 1e0..1e17 W/cm^2. PCHIP of monotone data is
    # monotone and interpolates its knots, so: non-decreasing everywhere; at each knot the
    # running maximum (hence >= rate_raw there); equal to rate_raw wherever rate_raw sets a
    # new maximum; well above rate_raw inside the cusps where rate_raw decreases (23 such
    # samples over 1e11.5..1e14.2, recorded by the reference's own validation).
    n0 = O.n_sellmeier(800e-9)
    kel = O.Keldysh(800e-9, 9.0, 0.64 * m_e, n0, monotone=True)
    I = np.logspace(11.5, 14.2, 20000)
    W = kel.rate(I)
    assert np.all(np.diff(W) >= 0.0)
    lI = np.linspace(0.0, 17.0, 6000)
    raw = np.maximum(kel.rate_raw(10.0 ** lI), 0.0)
    Wk = kel.rate(10.0 ** lI)
    sel = (lI >= 11.0) & (lI <= 14.5)
    run = np.array([raw[sel][:j + 1].max() for j in range(sel.sum())])
    assert np.allclose(Wk[sel], run, rtol=1e-10, atol=0)
    assert np.all(Wk[sel] >= raw[sel] * (1 - 1e-12))
    new_max = np.concatenate([[True], raw[sel][1:] > run[:-1]])
    assert np.allclose(Wk[sel][new_max], raw[sel][new_max], rtol=1e-10, atol=0)
    Ir = kel.rate_raw(I)
    dec = np.diff(np.log(np.maximum(Ir, 1e-300))) < 0
    assert dec.sum() > 0
    assert np.any(W[1:][dec] > Ir[1:][dec] * 1.01)
    # the default (monotone=False) is rate_raw itself
    assert np.array_equal(O.Keldysh(800e-9, 9.0, 0.64 * m_e, n0).rate(I), Ir)
