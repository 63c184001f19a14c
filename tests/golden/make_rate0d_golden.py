"""Write tests/golden/rate0d_ste.npz: input/output vectors of the reference's 0-D
re-integration (integre_0d, D3/sim/figures_article.py:168-194) for its STE branch.

The STE terms (the second population, its re-ionisation feedback on n_e and the
n_e + n_s <= rho_max rescaling) have no closed form that pins the oracle, so their
expected outputs come from the reference itself, run once in the build container:
`python tests/golden/make_rate0d_golden.py /root/reference`. The vectors are data
(inputs and outputs); no reference source is stored. Branches without STE are pinned by
closed forms in tests/test_oracle_postprocess.py and are not generated here.
"""
import os
import sys

import numpy as np


def main(ref_root):
    os.environ.setdefault("MPLBACKEND", "Agg")
    sys.path.insert(0, os.path.join(ref_root, ".github", "deliverable (3)", "sim"))
    import figures_article as fa  # noqa: E402  (the reference, read-only)

    prm = fa.Params()
    t_fs = np.arange(-150.0, 150.0, 0.5)
    rng = np.random.default_rng(7)
    peaks = np.array([3e13, 9e13, 4e14])
    I = peaks[:, None] * np.exp(-(t_fs / 60.0) ** 2)[None, :]
    I = I * (1.0 + 0.02 * rng.standard_normal(I.shape))
    out = {"t_fs": t_fs, "I": I}
    for name, (av, pg) in {"full": (True, True), "noav": (False, True), "notrap": (True, False)}.items():
        NE, NS = [], []
        for row in I:
            ne, ns = fa.integre_0d(t_fs, row, prm, avalanche=av, piegeage=pg, ste=True)
            NE.append(ne)
            NS.append(ns)
        out[f"ne_{name}"] = np.array(NE)
        out[f"ns_{name}"] = np.array(NS)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rate0d_ste.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
