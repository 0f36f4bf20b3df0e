"""The autotuner on the GPU: it measures, applies and caches a configuration,
and the configuration it picks changes speed only -- results stay bitwise
identical to those with the default knobs."""
import numpy as np
import pytest

from fastqsl_amd import synth

pytestmark = pytest.mark.gpu


def test_autotune_picks_knobs_that_do_not_change_results(tmp_path, monkeypatch):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fastqsl_amd as fq
    from fastqsl_amd import tune
    monkeypatch.setenv("FQSL_TUNE_CACHE", str(tmp_path / "tune.json"))
    f = synth.multipole(shape=(128, 128, 128), seed=5, device="cuda")
    g = np.linspace(0, 127, 512, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    s = [torch.as_tensor(a.ravel().copy(), device="cuda") for a in (X, Y, np.zeros_like(X))]
    before = tune.current()
    try:
        with fq.Field(*f) as field:
            tune.apply(tune.DEFAULTS)
            a = fq.trace_scott(field, *s, method="rkf45", step=0.25, tol=1e-6)
            torch.cuda.synchronize()
            a = {k: v.cpu().numpy() for k, v in a.items() if k != "_seeds"}
            cfg = tune.autotune(field, *s, method="rkf45", step=0.25, tol=1e-6, sample=1 << 17,
                                log=print)
            print("chosen", cfg)
            assert tune.current() == cfg
            assert set(cfg) == set(tune.SPACE)
            assert len(tune.load_cache()) == 1
            # every knob away from its default, results unchanged
            for alt in ({"refill_min": 1, "queues": 8, "grid_pct": 50}, {"order": 0}, {"order": 2, "order_buckets": 256},
                        {"scott_waves": 0, "refill_min": 32}, {"scott_waves": 2}, cfg):
                tune.apply(dict(tune.DEFAULTS, **alt))
                b = fq.trace_scott(field, *s, method="rkf45", step=0.25, tol=1e-6)
                torch.cuda.synchronize()
                for k in a:
                    assert np.array_equal(a[k], b[k].cpu().numpy(), equal_nan=True), (alt, k)
            # a cache hit applies the stored configuration without measuring
            tune.apply(tune.DEFAULTS)
            assert tune.autotune(field, *s, method="rkf45", step=0.25, tol=1e-6) == cfg
    finally:
        tune.apply(before)
