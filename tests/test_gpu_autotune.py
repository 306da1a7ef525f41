"""The autotuner: every candidate shape gives identical results (checked inside the tuner),
the choice is one of the candidates, and a tuned codec still matches the fixtures."""
import numpy as np
import pytest

import cases as C
import replay as R

pytestmark = pytest.mark.gpu


def test_tune_rsi16_and_apply(tmp_path, monkeypatch):
    monkeypatch.setenv("FFRS_TUNE_CACHE", str(tmp_path / "tune.json"))
    from libffrs_amd import autotune, capi
    ctx = capi.Context(0)
    res = autotune.tune_rsi16(ctx, 256, 32, count=2048)
    assert set(res["timings_ms"]) == {"1", "2", "4", "8", "16", "32"}
    assert res["g_encode"] in (1, 2, 4, 8, 16, 32) and res["g_repair"] in (1, 2, 4, 8, 16, 32)
    assert autotune.tune_rsi16(ctx, 256, 32, count=2048) == res  # from the cache
    codec = capi.RSi16Codec(ctx, 256, 32)
    autotune.apply_rsi16(codec)
    assert codec.groups() == (res["g_encode"], res["g_repair"])
    codec.close()
    r = autotune.tune_rs256(ctx, 16, count=4096)
    assert r["repair_mode"] in (1, 2)
    ctx.close()


@pytest.mark.parametrize("g", [2, 8])
def test_rsi16_fixtures_with_other_group_sizes(g):
    from backends import GpuBackend

    class Tuned(GpuBackend):
        def ri(self, n, ecc, vector=True):
            r = super().ri(n, ecc, vector)
            r._set_groups(min(g, ecc), min(g, ecc))
            return r
    be = Tuned()
    for name in ("n256_sub", "small_codes", "beyond_64"):
        R.check_rsi16_plain(be, name)
    R.check_rsi16_edges(be)
