"""Batched form: several mixed-precision instances of uneven Nr x Nt (and one with
STE on) advanced by grouped launches, checked against each instance's own
single-instance mixed kernels from the same input.

Bars:
* pass by pass, from identical states: the grouped plasma and GEMM passes equal
  the single-instance calls bit for bit (same device functions, same data); the
  grouped nonlinear pass agrees to FP32 rounding (rel-L2 <= 1e-6: the inlined body
  may contract multiply-adds differently);
* whole steps: rel-L2 <= 1e-5 against single runs after 5 steps;
* other GEMM tiles and persistent grids only change which block computes which
  output (every element keeps its K order): bit for bit equal to the default
  batched run; Nt/4 nonlinear threads change the FFT factorisation: rel-L2 <= 1e-5;
* a captured graph of the batched step reproduces the eager steps bit for bit on
  every replay.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tests._cases import MATERIAL, rel_l2  # noqa: E402

SHAPES = [  # (N, Nt, energy_uJ, enable_ste, w0, t_window)
    (129, 512, 2.0, False, 100e-6, 800e-15),
    (257, 1024, 6.0, False, 80e-6, 800e-15),
    (129, 1024, 20.0, False, 30e-6, 800e-15),
    (257, 512, 4.0, True, 90e-6, 800e-15),
    (129, 1024, 8.0, False, 60e-6, 400e-15),   # fine t grid: ~38 % of the columns masked
]


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _cfgs(nz=20, save_stride=5):
    import rapport_amd as R
    cfgs, envs = [], []
    for k, (N, Nt, E, ste, w0, tw) in enumerate(SHAPES):
        kw = dict(MATERIAL, nz=nz, Nt=Nt, N=N, R_factor=10.0, begin=0.0, end=nz * 10e-6,
                  t_window=tw, energy_uJ=E, enable_ste=ste, w0=w0, save_stride=save_stride,
                  rho_t_stride=0, write_output=False, verbose=False, precision="mixed")
        cfgs.append(R.Config(**kw))
        envs.append(R.envelope_lens_noise(lens_f=0.05 if k % 2 else None, noise=0.01, seed=10 + k))
    return cfgs, envs


def _singles(cfgs, envs, steps):
    import rapport_amd as R
    out = []
    for c, e in zip(cfgs, envs):
        it = R.Integrator(c, envelope=e)
        for _ in range(steps):
            it._pipeline_step(it.dz)
        out.append((it._mx.spec.clone(), it.rho.clone()))
        del it
    torch.cuda.synchronize()
    return out


def _state(b):
    return [(m._mx.spec.clone(), m.rho.clone()) for m in b.members]


def _same(x, y):
    """Equal values (IEEE: +0 == -0) at the same NaN positions. Masked frequency columns
    that the batched GEMM writes as +0 can be -0 in the single-instance product."""
    if x.is_complex():
        x, y = torch.view_as_real(x), torch.view_as_real(y)
    nx, ny = torch.isnan(x), torch.isnan(y)
    return torch.equal(nx, ny) and torch.equal(x[~nx], y[~ny])


def _full(m):
    mx = m._mx
    return dict(spec=mx.spec.clone(), w0=mx.w0.clone(), w1=mx.w1.clone(), rho=m.rho.clone(),
                rho_s=m.rho_s.clone(), slots=mx.slots.clone())


def _load(m, s):
    mx = m._mx
    for key, dst in (("spec", mx.spec), ("w0", mx.w0), ("w1", mx.w1), ("rho", m.rho),
                     ("rho_s", m.rho_s), ("slots", mx.slots)):
        dst.copy_(s[key])


def _single_pass(m, name):
    import ctypes
    from rapport_amd._lib import call, ptr, stream_ptr
    mx, mt = m._mx, m._mx.m
    ph = mt.phase_half(m.dz)
    st = stream_ptr()

    def g(B, C, a, rs, phase, bin_, cout, reset):
        call("fp_hankel_gemm_mixed", ptr(mt.Yh), ptr(mt.Yl), mt.y_scale_exp, ptr(B), ptr(C),
             m.Nr, m.Nt, ctypes.c_float(a), ptr(rs), ptr(phase), bin_, cout, reset, st)
    if name == "plasma":
        mx.plasma(m.rho, m.rho_s)
    elif name == "G1":
        g(mx.spec, mx.w1, mt.Rk, None, ph, mx._slot(0), mx._slot(1), None)
    elif name == "G2":
        g(mx.w1, mx.w0, mt.iRk, None, None, mx._slot(1), None, mx._slot(0))
    elif name == "nl":
        call("fp_nl_step_c64", ptr(mx.w0), ptr(mx.w0), ptr(m.rho), ptr(m._rho_s_arg()), 1, 1,
             ctypes.byref(m.nl.params(m.dz)), ctypes.byref(mt.tables), mx._slot(2), st)
    elif name == "G3":
        g(mx.w0, mx.w1, mt.Rk, None, ph, mx._slot(2), mx._slot(3), None)
    else:
        g(mx.w1, mx.spec, mt.iRk, mt.mask_r, None, mx._slot(3), mx._slot(0), None)


@pytest.fixture(scope="module")
def singles5():
    cfgs, envs = _cfgs()
    return _singles(cfgs, envs, 5)


def test_batch_passes_match_single_calls():
    from rapport_amd.batch import BatchedIntegrator
    _announce("batch passes vs single-instance calls from identical states, 2 steps")
    cfgs, envs = _cfgs()
    b = BatchedIntegrator(cfgs, envs)
    assert b.launches_per_step() == 3 + 3 + 4      # plasma: 512, 1024, 512+STE; nl same; 1 GEMM
    zc = [(d.zc_lo, d.zc_hi) for d in b._descs]
    assert zc[4][1] - zc[4][0] >= 256, zc           # whole masked tiles are skipped for instance 4
    for step in range(2):
        for name in ("plasma", "G1", "G2", "nl", "G3", "G4"):
            pre = [_full(m) for m in b.members]
            b.run_pass(name)
            torch.cuda.synchronize()
            post_b = [_full(m) for m in b.members]
            for k, m in enumerate(b.members):
                _load(m, pre[k])
                _single_pass(m, name)
            torch.cuda.synchronize()
            for k, m in enumerate(b.members):
                s1 = _full(m)
                for key in s1:
                    x, y = post_b[k][key], s1[key]
                    if name != "nl" or key not in ("w0", "slots"):
                        assert _same(x, y), f"step {step} {name} instance {k}: {key} differs"
                    elif key == "w0":
                        e = rel_l2(x.cpu().numpy(), y.cpu().numpy())
                        assert e <= 1e-6, f"step {step} nl instance {k}: w0 rel-L2 {e:.2e}"
                    else:
                        fx = x.cpu().view(torch.float32)[2].item()
                        fy = y.cpu().view(torch.float32)[2].item()
                        assert abs(fx - fy) <= 1e-6 * abs(fy)
                _load(m, post_b[k])
    assert all(int(m.t.status.item()) == 0 for m in b.members)


def test_batch_steps_match_single(singles5):
    from rapport_amd.batch import BatchedIntegrator
    _announce("batch default config vs single runs, 5 steps")
    cfgs, envs = _cfgs()
    b = BatchedIntegrator(cfgs, envs)
    for _ in range(5):
        b.step()
    torch.cuda.synchronize()
    assert max(float(m.rho.max()) for m in b.members) > 1e15   # plasma coupling exercised
    for k, ((s, r), (s1, r1)) in enumerate(zip(_state(b), singles5)):
        assert torch.isfinite(s).all()
        es = rel_l2(s.cpu().numpy(), s1.cpu().numpy())
        er = rel_l2(r.cpu().numpy(), r1.cpu().numpy())
        _announce(f"  instance {k}: spec rel-L2 {es:.2e}, rho rel-L2 {er:.2e}")
        assert es <= 1e-5 and er <= 1e-5


@pytest.fixture(scope="module")
def batch5():
    from rapport_amd.batch import BatchedIntegrator
    cfgs, envs = _cfgs()
    b = BatchedIntegrator(cfgs, envs)
    for _ in range(5):
        b.step()
    out = _state(b)
    b.close()
    return out


@pytest.mark.parametrize("tile", [(64, 64), (64, 128), (128, 64)])
def test_batch_tiles_and_persistent_grid_bitwise(batch5, tile):
    from rapport_amd.batch import BatchedIntegrator
    _announce(f"batch GEMM tile {tile} + persistent grids, 5 steps")
    cfgs, envs = _cfgs()
    b = BatchedIntegrator(cfgs, envs)
    for i in range(b.n):
        b.set_config(i, tile[0], tile[1], 0)
    b.set_grid("gemm", 37)
    b.set_grid("plasma", 19)
    b.set_grid("nl", 23)
    for _ in range(5):
        b.step()
    torch.cuda.synchronize()
    for k, ((s, r), (s1, r1)) in enumerate(zip(_state(b), batch5)):
        assert torch.equal(s, s1), f"instance {k}: spec differs, rel {rel_l2(s.cpu(), s1.cpu()):.2e}"
        assert torch.equal(r, r1), f"instance {k}: rho differs"


def test_batch_eight_wave_tile_bitwise(batch5):
    """256 x 128 tiles of 8 waves on the instances whose Nr allows them, the default tile on
    the others: the same per-element arithmetic, so the state equals the default run's."""
    from rapport_amd.batch import BatchedIntegrator
    _announce("batch GEMM tile 256 x 128 (8 waves), 5 steps")
    cfgs, envs = _cfgs()
    b = BatchedIntegrator(cfgs, envs)
    used = 0
    for i, m in enumerate(b.members):
        if m.Nr % 256 == 0:
            b.set_config(i, 256, 128, 0)
            used += 1
    assert used >= 1
    for _ in range(5):
        b.step()
    torch.cuda.synchronize()
    for k, ((s, r), (s1, r1)) in enumerate(zip(_state(b), batch5)):
        assert torch.equal(s, s1), f"instance {k}: spec differs, rel {rel_l2(s.cpu(), s1.cpu()):.2e}"
        assert torch.equal(r, r1), f"instance {k}: rho differs"
    b.close()


def test_batch_wide_nl_close(singles5):
    from rapport_amd.batch import BatchedIntegrator
    _announce("batch nonlinear Nt/4 threads, 5 steps")
    cfgs, envs = _cfgs()
    b = BatchedIntegrator(cfgs, envs)
    for i, m in enumerate(b.members):
        b.set_config(i, 0, 0, m.Nt // 4)
    for _ in range(5):
        b.step()
    torch.cuda.synchronize()
    for k, ((s, r), (s1, r1)) in enumerate(zip(_state(b), singles5)):
        e = rel_l2(s.cpu().numpy(), s1.cpu().numpy())
        _announce(f"  instance {k}: spec rel-L2 {e:.2e}")
        assert e < 1e-5
        assert rel_l2(r.cpu().numpy(), r1.cpu().numpy()) < 1e-4


def test_batch_graph_replay_matches_eager():
    from rapport_amd.batch import BatchedIntegrator
    _announce("batch graph replay vs eager, 3 steps")
    cfgs, envs = _cfgs()
    b = BatchedIntegrator(cfgs, envs)
    b.set_grid("gemm", 64)
    snap = b._snapshot()
    eager = []
    for _ in range(3):
        b.step()
        torch.cuda.synchronize()
        eager.append(_state(b))
    b._restore(snap)
    g = b.capture(1)
    for j in range(3):
        g.replay()
        torch.cuda.synchronize()
        _ = sum(float(m.rho.sum()) for m in b.members)    # eager work between replays
        for k, ((s, r), (se, re)) in enumerate(zip(_state(b), eager[j])):
            assert torch.equal(s, se), f"replay {j} instance {k}: spec differs"
            assert torch.equal(r, re), f"replay {j} instance {k}: rho differs"


def test_autotune_and_propagate_match_singles():
    import rapport_amd as R
    from rapport_amd.batch import BatchedIntegrator
    _announce("autotune + batched propagate (graph) vs single propagate, 11 steps")
    cfgs, envs = _cfgs(nz=10, save_stride=5)
    b = BatchedIntegrator(cfgs, envs, tune=True)
    assert set(b.config) >= {"gemm", "nl", "grid"}
    res = b.propagate(graph=True)
    for k, (c, e) in enumerate(zip(cfgs, envs)):
        r1 = R.Integrator(c, envelope=e).propagate()
        for key in ("fluence_rz", "rho_rz", "Imax_z", "z"):
            err = rel_l2(res[k][key], r1[key])
            assert err < 1e-5, f"instance {k} {key}: rel-L2 {err:.2e}"
        assert np.array_equal(res[k]["z"], r1["z"])
