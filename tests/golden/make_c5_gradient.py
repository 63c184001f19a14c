"""Write tests/golden/c5_gradient.npz: the FP64 oracle's gradient of the fluence
mismatch loss with respect to 32 Laguerre-Gauss input coefficients, after 1000
z-steps at Nr x Nt = 512 x 2048 (dz = 10 um, 3 uJ, 85 fs, w0 = 100 um, STE off).

The oracle is the NumPy FP64 restatement of the reference (oracle/reference_numpy.py)
with its FP64 discrete adjoint (oracle/adjoint_numpy.py, pinned by central finite
differences in tests/test_oracle_adjoint.py). Nothing of the reference is imported.

Problem (the same as bench.py's adjoint workload): P = 16 modes, hidden coefficients
and start point from rapport_amd.adjoint.lg_problem(16, seed=5); the gradient is taken
at theta = theta0 + (hidden - theta0) / 2 against the target F* = fluence of the FP64
forward run from the hidden coefficients.

Run once on the CPU (about 1.5-2 h on 8 cores):

    python tests/golden/make_c5_gradient.py [--workdir DIR]

The forward states of the theta run are kept in a memory-mapped file under --workdir
(16 GiB, deleted at the end); the target run proceeds in a second process meanwhile.
Progress is saved as it goes, so an interrupted run resumes where it stopped.
"""
import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

WORKLOAD = dict(nz=999, Nt=2048, N=513, R_factor=15.0, begin=0.0, end=999 * 10e-6,
                energy_uJ=3.0, wavelength=800e-9, n2=3.2e-20, Ui_eV=9.0, rho_max=2.1e22,
                tau_c=1.27e-15, tau_r=150e-15, delta_t=85e-15, w0=100e-6, enable_ste=False,
                save_stride=10 ** 6, rho_t_stride=0, write_output=False, verbose=False)
P, SEED = 16, 5


def lg_problem(P, seed, amp_sigma=0.1, phase_sigma=0.3):
    """Same draw as rapport_amd.adjoint.lg_problem (kept here so the script needs no GPU
    package import)."""
    rng = np.random.default_rng(seed)
    damp = 1.0 / (1.0 + np.arange(P))
    theta0 = np.zeros(2 * P)
    theta0[0] = 1.0
    hidden = theta0.copy()
    hidden[:P] += amp_sigma * rng.standard_normal(P) * damp
    hidden[P:] += phase_sigma * rng.standard_normal(P) * damp
    return theta0, hidden


def _setup():
    from oracle import adjoint_numpy as A
    from oracle import reference_numpy as O
    it = O.Integrator(O.Config(**WORKLOAD))
    return A, A.Adjoint(it)


def _save(path, arr):
    np.save(path + ".tmp.npy", arr)
    os.replace(path + ".tmp.npy", path)


def _target(hidden, n, wd):
    """Fluence of the FP64 run from the hidden coefficients (resumable: the state is
    saved every 50 steps under wd)."""
    done = os.path.join(wd, "c5_target_F.npy")
    if os.path.exists(done):
        return
    A, oa = _setup()
    ck, kf = os.path.join(wd, "c5_target_u.npy"), os.path.join(wd, "c5_target_k.npy")
    if os.path.exists(ck) and os.path.exists(kf):
        u, k0 = np.load(ck), int(np.load(kf))
    else:
        u, k0 = A.lg_field(hidden, oa.g, oa.cfg, P), 0
    for k in range(k0, n):
        u = oa.step(u)
        if k % 50 == 49:
            _save(ck, u)
            _save(kf, np.array(k + 1))
            print(f"[target] step {k + 1}/{n}", flush=True)
    _save(done, A.fluence(u, oa.g))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workdir", default=os.path.join(ROOT, "_work"))
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "c5_gradient.npz"))
    ap.add_argument("--steps", type=int, default=WORKLOAD["nz"] + 1)
    args = ap.parse_args()
    n = args.steps
    wd = os.path.join(args.workdir, f"c5_{n}")
    os.makedirs(wd, exist_ok=True)
    theta0, hidden = lg_problem(P, SEED)
    theta = theta0 + 0.5 * (hidden - theta0)
    t0 = time.time()
    pt = mp.get_context("fork").Process(target=_target, args=(hidden, n, wd))
    pt.start()
    A, oa = _setup()
    # forward run from theta: every state kept in a memory-mapped file (resumable)
    path = os.path.join(wd, "c5_states.npy")
    kf = os.path.join(wd, "c5_fwd_k.npy")
    u = A.lg_field(theta, oa.g, oa.cfg, P)
    k0 = int(np.load(kf)) if os.path.exists(kf) and os.path.exists(path) else 0
    states = np.lib.format.open_memmap(path, mode="r+" if k0 else "w+", dtype=np.complex128,
                                       shape=(n + 1,) + u.shape)
    if k0:
        u = np.array(states[k0])
    else:
        states[0] = u
    for k in range(k0, n):
        u = oa.step(u)
        states[k + 1] = u
        if k % 10 == 9 or k == n - 1:
            states.flush()
            _save(kf, np.array(k + 1))
        if k % 100 == 99:
            print(f"[forward] step {k + 1}/{n} ({time.time() - t0:.0f} s)", flush=True)
    pt.join()
    F_t = np.load(os.path.join(wd, "c5_target_F.npy"))
    L, g = A.fluence_loss(np.array(states[n]), oa.g, F_t)
    # adjoint sweep (resumable: the running gradient is saved every 10 steps)
    gf, gkf = os.path.join(wd, "c5_adj_g.npy"), os.path.join(wd, "c5_adj_k.npy")
    ks = n - 1
    if os.path.exists(gf) and os.path.exists(gkf):
        g, ks = np.load(gf), int(np.load(gkf))
    for k in range(ks, -1, -1):
        g = oa.step_adj(np.array(states[k]), g)
        if k % 10 == 0:
            _save(gf, g)
            _save(gkf, np.array(k - 1))
        if k % 100 == 0:
            print(f"[adjoint] step {k}/{n} ({time.time() - t0:.0f} s)", flush=True)
    grad = A.lg_field_grad(theta, oa.g, oa.cfg, P, g)
    del states
    np.savez(args.out, P=P, seed=SEED, steps=n, theta=theta, hidden=hidden, F_target=F_t,
             loss=L, grad=grad)
    for f in os.listdir(wd):
        os.remove(os.path.join(wd, f))
    os.rmdir(wd)
    print(f"[done] loss {L:.6e} |grad| {np.linalg.norm(grad):.6e} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
