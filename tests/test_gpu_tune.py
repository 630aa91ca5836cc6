"""The autotuner on the GPU at small sizes: every stage measures, the stored
result is in range and restores the engine, and two-site updates with the
tuned settings reconstruct theta as the default settings do."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tuner_runs_small(tmp_path):
    from mpsim_amd import _native, tune
    _native.load()  # fail loudly if the library is missing
    import torch
    from mpsim_amd.engine import Engine, SvdConfig
    eng = Engine.get()
    before = (eng.split_ways, eng.split_min, eng.split_rows)
    path = str(tmp_path / "tuned.json")
    out = tune.tune(eng, chis=(8, 32, 64), reps=1, path=path, conc_chis=(64,), layer=8,
                    batch_sizes=(2, 8))
    assert out["block_switch"] >= 16 and out["inner_sweeps"] in (1, 2) and out["method"] in (0, 1)
    assert 1 <= out["split_ways"] <= 4 and out["split_min"] >= 1 and out["split_rows"] >= 1
    assert (eng.split_ways, eng.split_min, eng.split_rows) == before
    times = [v for r in out["measurements"] for k, v in r.items() if k.startswith("us_per_update")]
    assert times and all(np.isfinite(t) and t > 0 for t in times)
    stages = {r["stage"] for r in out["measurements"]}
    assert {"pivot", "method", "split"} <= stages
    # split_min is measured only when splitting the layer paid at all
    assert ("split_min" in stages) == (out["split_ways"] > 1)
    cfg = tune.tuned_config(path)
    items = tune._problems(eng, 64, 6, seed=5)
    ref = eng.two_site(items, SvdConfig())
    got = eng.two_site(items, cfg)
    torch.cuda.synchronize()
    for r, g in zip(ref, got):
        rr = torch.einsum("lak,kbr->labr", r.a_out, r.b_out).cpu().numpy()
        gg = torch.einsum("lak,kbr->labr", g.a_out, g.b_out).cpu().numpy()
        assert np.max(np.abs(rr - gg)) <= 1e-5 * np.max(np.abs(rr))
