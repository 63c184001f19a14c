"""Batched form at the sweep's own shapes: one instance of every (Nr, Nt) class that the
64-instance sweep (random_sweep(64, seed=0): Nr 256-1024 x Nt 1024-4096, 0.5-20 uJ,
40-200 fs, w0 50-150 um) contains, so every batched kernel instantiation the sweep uses
runs here (carrier scan and nonlinear step at Nt = 1024, 2048, 4096; the GEMM tiles the
tuner picks, 256 x 128 included).

Bars:
* tuned + graph-captured batched steps against each instance's own single-instance mixed
  run from the same input, 3 steps: rel-L2 <= 1e-5 on the spectral state and the densities
  (the grouped nonlinear pass may contract multiply-adds differently: FP32 rounding);
* the same with a forced configuration (256 x 128 tiles where Nr allows, 128 x 128
  elsewhere; Nt/4 nonlinear threads, i.e. the other FFT factorisation; persistent grids):
  rel-L2 <= 1e-5;
* the two smallest classes against the FP64 oracle after 3 steps, at the mixed path's
  tolerance: field, fluence and electron density rel-L2 <= 1e-4.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tests._cases import rel_l2  # noqa: E402

STEPS = 3


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _classes():
    """First instance of each (Nr, Nt) class of the sweep, with nz = STEPS - 1."""
    from dataclasses import replace
    from rapport_amd.batch import random_sweep
    cfgs, envs = random_sweep(64, seed=0)
    seen, out_c, out_e, seeds = set(), [], [], []
    for k, (c, e) in enumerate(zip(cfgs, envs)):
        key = (c.N - 1, c.Nt)
        if key in seen:
            continue
        seen.add(key)
        out_c.append(replace(c, nz=STEPS - 1, end=c.begin + (STEPS - 1) * c.dz,
                             save_stride=10 ** 6))
        out_e.append(e)
        seeds.append(1000 + k)            # random_sweep's noise seed of instance k
    order = np.argsort([c.N * c.Nt for c in out_c], kind="stable")
    _classes.seeds = [seeds[i] for i in order]
    return [out_c[i] for i in order], [out_e[i] for i in order]


@pytest.fixture(scope="module")
def case():
    import rapport_amd as R
    cfgs, envs = _classes()
    shapes = sorted({(c.N - 1, c.Nt) for c in cfgs})
    assert {s[1] for s in shapes} == {1024, 2048, 4096} and {s[0] for s in shapes} >= {256, 1024}
    singles = []
    for c, e in zip(cfgs, envs):
        it = R.Integrator(c, envelope=e)
        for _ in range(STEPS):
            it._pipeline_step(it.dz)
        torch.cuda.synchronize()
        singles.append((it._mx.spec.clone(), it.rho.clone()))
        del it
    return cfgs, envs, singles


def _check(b, singles, label):
    torch.cuda.synchronize()
    worst = 0.0
    for k, (m, (s1, r1)) in enumerate(zip(b.members, singles)):
        s, r = m._mx.spec, m.rho
        assert torch.isfinite(s).all(), (label, k)
        es = rel_l2(s.cpu().numpy(), s1.cpu().numpy())
        er = rel_l2(r.cpu().numpy(), r1.cpu().numpy())
        worst = max(worst, es, er)
        assert es <= 1e-5 and er <= 1e-5, f"{label}: instance {k} ({m.Nr} x {m.Nt}) " \
                                           f"spec {es:.2e} rho {er:.2e}"
    _announce(f"  {label}: worst rel-L2 {worst:.2e} over {len(singles)} instances")


def test_sweep_classes_tuned_graph_vs_single(case):
    from rapport_amd.batch import BatchedIntegrator
    cfgs, envs, singles = case
    _announce(f"sweep classes {[(c.N - 1, c.Nt) for c in cfgs]}: tuned + graph, {STEPS} steps")
    b = BatchedIntegrator(cfgs, envs, tune=True)
    _announce(f"  tuned: {dict((k, v) for k, v in b.config.items() if k != 'timings_ms')}")
    # the tuned per-class shapes in grouped launches, whichever launch mode the tuner chose
    # (one launch per pass has its own test below)
    b.set_merged(False)
    g = b.capture(1)
    for _ in range(STEPS):
        g.replay()
    _check(b, singles, "tuned + graph")
    assert max(float(m.rho.max()) for m in b.members) > 1e14      # plasma coupling exercised
    b.close()


def test_sweep_classes_forced_config_vs_single(case):
    from rapport_amd.batch import BatchedIntegrator
    cfgs, envs, singles = case
    _announce(f"sweep classes: 256 x 128 / 128 x 128 tiles, Nt/4 threads, persistent grids")
    b = BatchedIntegrator(cfgs, envs)
    for i, m in enumerate(b.members):
        if m.Nr % 256 == 0:
            b.set_config(i, 256, 128, m.Nt // 4)
        else:
            b.set_config(i, 128, 128, m.Nt // 4)
    b.set_grid("gemm", 512)
    b.set_grid("plasma", 256)
    b.set_grid("nl", 256)
    for _ in range(STEPS):
        b.step()
    _check(b, singles, "forced config")
    b.close()


def test_sweep_smallest_classes_vs_fp64_oracle(case):
    from oracle import reference_numpy as O
    cfgs, envs, _ = case
    for c, e, seed in list(zip(cfgs, envs, _classes.seeds))[:2]:
        _announce(f"sweep class {c.N - 1} x {c.Nt} ({c.energy_uJ:.2f} uJ): batched mixed vs "
                  f"FP64 oracle, {STEPS} steps")
        from rapport_amd.batch import BatchedIntegrator
        b = BatchedIntegrator([c], [e])
        res = b.propagate(graph=True)[0]
        u = b.members[0].u.cpu().numpy()
        from dataclasses import fields
        kw = {f.name: getattr(c, f.name) for f in fields(c) if f.init and f.name != "precision"}
        oi = O.Integrator(O.Config(**kw),
                          envelope=O.envelope_lens_noise(lens_f=None, noise=0.01, seed=seed))
        ro = oi.propagate()
        eu = rel_l2(u, oi.u)
        ef = rel_l2(res["fluence_rz"], ro["fluence_rz"])
        er = rel_l2(res["rho_rz"], ro["rho_rz"])
        _announce(f"  field {eu:.2e} fluence {ef:.2e} rho {er:.2e}")
        assert eu <= 1e-4 and ef <= 1e-4 and er <= 1e-4
        b.close()


def test_sweep_classes_merged_single_launch_per_pass_vs_single(case):
    from rapport_amd.batch import BatchedIntegrator
    cfgs, envs, singles = case
    _announce("sweep classes: one launch per pass (merged work lists), graph, 3 steps")
    b = BatchedIntegrator(cfgs, envs)
    b.set_merged(True)
    assert b.launches_per_step() == 6
    g = b.capture(1)
    for _ in range(STEPS):
        g.replay()
    _check(b, singles, "merged")
    b.close()


def test_sweep_classes_mixed_launch_modes_vs_single(case):
    """G1 and G3 as one launch each over the merged work list, the other passes grouped."""
    from rapport_amd.batch import BatchedIntegrator
    cfgs, envs, singles = case
    _announce("sweep classes: G1 / G3 merged, other passes grouped, graph, 3 steps")
    b = BatchedIntegrator(cfgs, envs)
    grouped = b.launches_per_step()
    b.set_merged_passes((1 << 1) | (1 << 4))
    assert 6 <= b.launches_per_step() <= grouped
    g = b.capture(1)
    for _ in range(STEPS):
        g.replay()
    _check(b, singles, "G1 / G3 merged")
    b.close()
