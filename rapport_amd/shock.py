"""1-D SPM + optical-shock model on the GPU.

Drop-in for the numerical core of ``D3/sim/self_steepening.py`` (D3 = ``.github/deliverable
(3)``): ``propagate_shock`` (``:80-105``) integrates, by classical RK4 in z,

    dE/dz = i K0 n2 |E|^2 E - (n2/c) d(|E|^2 E)/dt,

with the t-derivative taken spectrally (``ifft(i w fft(.))``, ``:60-62``). Here the whole
z march of a pulse is one workgroup of ``fp_shock_propagate`` (field in registers, the two
transforms of each right-hand side in LDS, FP64), and ``propagate_shock_batch`` runs many
pulses -- e.g. the pulse-duration scan of ``fig_shock_scaling`` (``:296-315``) -- in one
launch. The closed form ``shock_distance_theory`` and the slope diagnostics are host
NumPy, as in the reference.
"""
from __future__ import annotations

import numpy as np
import torch
from scipy.constants import c as C_LIGHT

from ._lib import call, ptr, stream_ptr

SLOPE_THRESHOLD = 3.0
THRESHOLD_FRACTION = 1.0 - 1.0 / SLOPE_THRESHOLD
SIZES = (1024, 2048, 4096, 8192)


def shock_distance_theory(t_p_s, I_p, n2=3.54e-20):
    """z_shock = c t_p / (3 sqrt(2) e^{-1/2} n2 I_p) for a Gaussian I_p exp(-t^2/t_p^2)
    (self_steepening.py:108-116): the steepest slope of the Gaussian in the Burgers
    breaking condition."""
    coeff = 1.0 / (3.0 * np.sqrt(2.0) * np.exp(-0.5))
    return coeff * C_LIGHT * t_p_s / (n2 * I_p)


def save_plan(nz, nsave=None):
    """(stride, n_saves) of the reference's saving rule: nsave defaults to min(nz, 240),
    a plane is kept after every `stride` = max(1, nz // nsave) steps, plus the input."""
    nsave = nsave or min(nz, 240)
    stride = max(1, nz // nsave)
    return stride, 1 + nz // stride


def propagate_shock_batch(t_list, E0_rows, z_max, nz, lam0=800e-9, n0=1.4533, n2=3.54e-20,
                          shock=True, nsave=None, device=None):
    """Propagate every pulse of `E0_rows` (n_rows x N, sqrt(W/m^2)) over its own time grid
    (`t_list`: one array per row, or one shared array) and distance (`z_max`: scalar or one
    per row) in `nz` RK4 steps. `shock` may be a bool or one bool per row.

    Returns (z, E): z is n_rows x n_saves (m), E is n_rows x n_saves x N complex128 (a
    torch tensor on the device); plane 0 is the input."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.type != "cuda":
        raise ValueError("propagate_shock_batch runs on the HIP device")
    if torch.is_tensor(E0_rows):
        E0 = E0_rows.to(device=dev, dtype=torch.complex128)
    else:
        E0 = torch.as_tensor(np.asarray(E0_rows), dtype=torch.complex128, device=dev)
    if E0.dim() == 1:
        E0 = E0[None, :]
    E0 = E0.contiguous()
    n_rows, N = E0.shape
    if N not in SIZES:
        raise ValueError(f"N = {N}: supported sizes are {SIZES}")
    if nz < 0:
        raise ValueError("nz must be >= 0")
    t_rows = [np.asarray(t, dtype=np.float64) for t in (
        t_list if (len(t_list) == n_rows and np.ndim(t_list[0]) == 1) else [t_list] * n_rows)]
    zmax = np.broadcast_to(np.asarray(z_max, dtype=np.float64), (n_rows,))
    sk = np.broadcast_to(np.asarray(shock, dtype=bool), (n_rows,))
    K0 = 2 * np.pi / lam0
    rows = np.empty((n_rows, 4), dtype=np.float64)
    dzs = np.empty(n_rows)
    for i in range(n_rows):
        if t_rows[i].size != N:
            raise ValueError(f"row {i}: time grid has {t_rows[i].size} samples, the pulse {N}")
        dt = t_rows[i][1] - t_rows[i][0]
        dzs[i] = zmax[i] / nz if nz else 0.0
        rows[i] = (dt, dzs[i], K0 * n2, (n2 / C_LIGHT) if sk[i] else 0.0)
    stride, n_saves = save_plan(nz, nsave)
    out = torch.empty((n_rows, n_saves, N), dtype=torch.complex128, device=dev)
    work = torch.empty((n_rows, 3, N), dtype=torch.complex128, device=dev) if N == 8192 else None
    m = np.arange(N)
    tw = torch.as_tensor(np.exp(-2j * np.pi * m / N), device=dev)
    rows_d = torch.as_tensor(rows, device=dev)
    with torch.cuda.device(dev):
        call("fp_shock_propagate", ptr(torch.view_as_real(E0)), ptr(torch.view_as_real(out)),
             ptr(torch.view_as_real(work)) if work is not None else None, ptr(rows_d), n_rows, N,
             int(nz), int(stride), ptr(torch.view_as_real(tw)), stream_ptr())
    steps = [0] + [i + 1 for i in range(nz) if (i + 1) % stride == 0]
    z = np.array([[s * dzs[r] for s in steps] for r in range(n_rows)])
    return z, out


def propagate_shock(t_s, E0, z_max, nz, lam0=800e-9, n0=1.4533, n2=3.54e-20, shock=True,
                    nsave=None):
    """Same signature and result as the reference: (z, E_saved) as NumPy arrays, E_saved of
    shape (n_saves, len(t_s))."""
    z, E = propagate_shock_batch([np.asarray(t_s)], np.asarray(E0)[None, :], z_max, nz, lam0, n0,
                                 n2, shock, nsave)
    return z[0], E[0].cpu().numpy()


def max_slope_vs_z(Es, dt):
    """max_t |dI/dt| of every saved plane (self_steepening.py:263-266)."""
    I = np.abs(Es) ** 2
    return np.max(np.abs(np.gradient(I, dt, axis=-1)), axis=-1)


def measured_shock_distance(z, Es, dt, threshold=SLOPE_THRESHOLD):
    """z where max|dI/dt| first reaches `threshold` x its input value, linearly interpolated
    between saved planes (self_steepening.py:269-293); NaN if never reached. Along a Burgers
    characteristic |dI/dt| grows as 1 / (1 - z / z_shock), so the crossing sits at
    (1 - 1/threshold) z_shock."""
    slope = max_slope_vs_z(Es, dt)
    target = threshold * slope[0]
    idx = int(np.argmax(slope >= target))
    if slope[idx] < target:
        return np.nan
    if idx == 0:
        return z[0]
    s0, s1 = slope[idx - 1], slope[idx]
    return z[idx - 1] + (z[idx] - z[idx - 1]) * (target - s0) / (s1 - s0)
