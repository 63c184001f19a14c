"""Host-side checks of the batched form (no GPU): the sweep generator and the
validation that fp_batch_plan_create does before it touches the device."""
import ctypes

import numpy as np
import pytest


def test_random_sweep_ranges_and_determinism():
    from rapport_amd.batch import random_sweep
    cfgs, envs = random_sweep(64, seed=3)
    cfgs2, _ = random_sweep(64, seed=3)
    assert len(cfgs) == len(envs) == 64
    assert [c.N for c in cfgs] == [c.N for c in cfgs2]
    assert {c.N - 1 for c in cfgs} <= {256, 512, 768, 1024}
    assert {c.Nt for c in cfgs} <= {1024, 2048, 4096}
    assert len({c.N for c in cfgs}) == 4 and len({c.Nt for c in cfgs}) == 3
    E = np.array([c.energy_uJ for c in cfgs])
    assert E.min() >= 0.5 and E.max() <= 20.0
    # log-uniform: about half of the draws below the geometric mean sqrt(0.5 * 20) = 3.16
    assert 16 <= int(np.sum(E < np.sqrt(10.0))) <= 48
    fw = np.array([c.delta_t for c in cfgs]) * 1e15
    w0 = np.array([c.w0 for c in cfgs]) * 1e6
    assert fw.min() >= 40 and fw.max() <= 200 and w0.min() >= 50 and w0.max() <= 150
    for c in cfgs:
        assert c.precision == "mixed" and c.nz == 1999
        assert abs(c.dz - 10e-6) < 1e-15


def _desc(**over):
    from rapport_amd import _lib
    d = _lib.BatchInstance()
    d.Nr, d.Nt = 256, 1024
    d.pp.Nr, d.pp.Nt, d.pp.table_size = 256, 1024, 4096
    d.np.Nr, d.np.Nt = 256, 1024
    for f in ("Yh", "Yl", "spec", "w0", "w1", "rho", "phase", "mask_r", "slots", "W_lut", "tw"):
        setattr(d, f, 0x1000)
    for f in ("tw", "mA", "mB", "mC", "Rf", "knots", "inv_h", "pcoef_s", "rw"):
        setattr(d.nt, f, 0x1000)
    for k, v in over.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("over,msg", [
    (dict(Nt=300), "Nt"),
    (dict(Nr=100), "Nr"),
    (dict(spec=None), "null pointer"),
    (dict(enable_ste=1), "STE"),
])
def test_plan_create_rejects_bad_descriptors(over, msg):
    from rapport_amd import _lib
    good = _desc()
    bad = _desc(**over)
    arr = (_lib.BatchInstance * 2)(good, bad)
    plan = ctypes.c_void_p()
    with pytest.raises(RuntimeError, match=msg):
        _lib.call("fp_batch_plan_create", arr, 2, ctypes.byref(plan))
    assert not plan.value
    with pytest.raises(RuntimeError, match="instance 1"):
        _lib.call("fp_batch_plan_create", arr, 2, ctypes.byref(plan))


def test_plan_calls_reject_null_plan():
    from rapport_amd import _lib
    for name, args in (("fp_batch_step_mixed", (None, None)), ("fp_batch_pass", (None, 0, None)),
                       ("fp_batch_set_grid", (None, 0, 0)),
                       ("fp_batch_set_config", (None, 0, 0, 0, 0))):
        with pytest.raises(RuntimeError):
            _lib.call(name, *args)
