"""Launch configurations and the autotuner (GPU tests): every tile height and
tile order gives the same traces and images bit for bit, and the autotuner
picks one, caches it per signature and returns an equivalent batch."""
import numpy as np
import pytest

from tests.problems import Synthetic
from tests.test_gpu_graph import _batch, _gradient, _same, _uneven

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tile_heights_agree_bitwise(gpu, dtype):
    from pysit_amd.device import DeviceBatch
    probs = _uneven()
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    outs = []
    for tx in (8, 4, 2):
        p0 = probs[0]
        b = DeviceBatch([p.geometry() for p in probs], p0.model, p0.deltas, p0.dt, p0.order, dtype=dtype,
                        tile_x=tx, graph_steps=0)
        assert b.tile_x == tx
        outs.append(_gradient(b, res))
    assert _same(outs[0][0], outs[1][0]) and _same(outs[0][1], outs[1][1])


def test_tile_orders_and_heights_agree_for_alike_shots(gpu):
    from pysit_amd.device import DeviceBatch
    probs = [Synthetic((60, 75), ((6, 7), (5, 8)), (0.01, 0.012), 90, 8, seed=s, n_src=1 + s) for s in range(3)]
    for p in probs[1:]:
        p.model, p.Cpad, p.dt = probs[0].model, probs[0].Cpad, probs[0].dt
    rng = np.random.default_rng(2)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    ref = None
    for tx in (8, 4, 2):
        for order in ("shot", "interleave"):
            p0 = probs[0]
            b = DeviceBatch([p.geometry() for p in probs], p0.model, p0.deltas, p0.dt, 8, dtype=np.float32,
                            tile_x=tx, tile_order=order)
            out = _gradient(b, res)
            if ref is None:
                ref = out
            assert _same(out[0], ref[0]) and _same(out[1], ref[1]), (tx, order)


def test_unsupported_tile_height_is_rejected(gpu):
    from pysit_amd.device import DeviceBatch
    p = Synthetic((40, 50), ((4, 4), (4, 4)), (0.01, 0.01), 20, 4, seed=1)
    with pytest.raises(ValueError, match="tile_x"):
        DeviceBatch([p.geometry()], p.model, p.deltas, p.dt, 4, dtype=np.float32, tile_x=6)


def test_autotuner_picks_caches_and_matches(gpu, tmp_path, monkeypatch):
    from pysit_amd import autotune
    from pysit_amd.device import DeviceBatch
    monkeypatch.setenv("PYSIT_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.clear_cache()
    probs = _uneven()
    p0 = probs[0]
    geos = [p.geometry() for p in probs]
    report = []
    b = autotune.tuned_batch(geos, p0.model, p0.deltas, p0.dt, 8, dtype=np.float32, report=report)
    print(report, flush=True)
    assert len(report) >= 3 and all(t > 0 for _, t in report)
    cfg = {"tile_x": b.tile_x, "tile_order": b.tile_order, "graph_steps": b.graph_steps}
    assert cfg in [c for c, _ in report]
    # cached in memory and in the file: no timing the second time
    again = []
    b2 = autotune.tuned_batch(geos, p0.model, p0.deltas, p0.dt, 8, dtype=np.float32, report=again)
    assert again == [] and b2.tile_x == b.tile_x and b2.graph_steps == b.graph_steps
    autotune.clear_cache()
    b3 = autotune.tuned_batch(geos, p0.model, p0.deltas, p0.dt, 8, dtype=np.float32, report=again)
    assert again == [] and b3.tile_x == b.tile_x
    rng = np.random.default_rng(9)
    res = [rng.standard_normal((p.nt, p.n_rec)) for p in probs]
    ref = _gradient(DeviceBatch(geos, p0.model, p0.deltas, p0.dt, 8, dtype=np.float32), res)
    out = _gradient(b, res)
    assert _same(out[0], ref[0]) and _same(out[1], ref[1])
