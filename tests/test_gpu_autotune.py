"""The autotuner (bem_amd/autotune.py) on the post-tip workload shape:
every candidate runs, the winner is applied, and the tuned configuration
gives the same per-group attempt/accept counts as the default one (launch
shape and schedule change only the order of the field's partial sums;
state agrees to 1e-11)."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def test_autotune_picks_a_candidate_and_keeps_results(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bem_amd import SpecificParticleTracer, autotune
    from bem_amd.bem.geometry import PostTipBEMGeometry
    from bem_amd.emission import cap_source
    from bem_amd.particles import ParticleGroup

    d = golden("bem_post400.npz")
    R, H = float(d["R"]), float(d["H"])
    geo = PostTipBEMGeometry(-1e8, R, H, n_segments=400, sigma=d["sigma"])
    n = 512
    src = cap_source(n, R, H, seed=3)
    pg = ParticleGroup(dict(src, weight=np.full(n, 1.602176634e-19), status=np.ones(n), species="electron"))
    tr = SpecificParticleTracer(pg, 16, geo, [90e-6], rtol=1e-9, z_max=100e-6, t_max=2e-13)
    ref = tr.trace()
    res = autotune.tune(tr, attempts=4, log=_say)
    assert len(res["table"]) >= 6 and res["best"] in res["table"]
    assert all(r["ms"] > 0 for r in res["table"])
    _say(f"autotune best: {res['best']}")
    out = tr.trace()
    np.testing.assert_array_equal(out["attempts"], ref["attempts"])
    np.testing.assert_array_equal(out["accepted"], ref["accepted"])
    np.testing.assert_allclose(out["pos"], ref["pos"], rtol=1e-11, atol=1e-11 * np.abs(ref["pos"]).max())
