"""Mixed-size plans (one persistent work-queue launch per field, eager and replayed from a
HIP graph) against the golden fixtures written from the reference's own code: the mixed
sets, and every plain set run as a one-class plan."""
import pytest

import cases as C
import replay as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plan():
    from backends import PlanBackend
    return PlanBackend()


def test_rs256_mixed_plan(plan):
    R.check_rs256_mixed(plan)


def test_rsi16_mixed_plan(plan):
    R.check_rsi16_mixed(plan)


@pytest.mark.parametrize("name", C.RS256_PLAIN_CASES)
def test_rs256_plain_as_plan(plan, name):
    R.check_rs256_plain(plan, name)


@pytest.mark.parametrize("shape", C.RS256_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rs256_shapes_as_plan(plan, shape):
    R.check_rs256_shape(plan, *shape)


@pytest.mark.parametrize("name", C.RSI16_PLAIN_CASES)
def test_rsi16_plain_as_plan(plan, name):
    R.check_rsi16_plain(plan, name)


@pytest.mark.parametrize("shape", C.RS256_REF_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rs256_reference_test_shapes_as_plan(plan, shape):
    R.check_rs256_reftest(plan, *shape)


@pytest.mark.parametrize("shape", C.RSI16_REF_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rsi16_reference_test_shapes_as_plan(plan, shape):
    """Includes block_len 32768 and 65536 and ecc_len 2048 (the LDS kernels) in plans."""
    R.check_rsi16_reftest(plan, *shape)


@pytest.mark.parametrize("shape", C.RSI16_DIV_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rsi16_simd_word_divergence_as_plan(plan, shape):
    R.check_rsi16_divergence(plan, *shape)


def _rs256_all(mode):
    def f(plan):
        for e in plan.classes():
            plan.set_class_mode(e, mode)
    return f


def _rsi16_all(l_encode, g_repair):
    def f(plan):
        for n, e in plan.classes():
            g = g_repair if g_repair <= min(64, e) and (64 // g_repair) * 9 * e * 4 <= 160 * 1024 else 0
            plan.set_class_groups(n, e, l_encode if e <= 32 else 0, g)
    return f


@pytest.mark.parametrize("mode", [1, 2])
def test_rs256_mixed_plan_every_class_mode(mode):
    """Every class one codeword per lane (1) or per wave (2): the mixed fixtures unchanged."""
    from backends import PlanBackend
    R.check_rs256_mixed(PlanBackend(graph_replays=2, shape=_rs256_all(mode)))


@pytest.mark.parametrize("shape", [(1, 8), (4, 4), (16, 2)], ids=lambda s: f"l{s[0]}_g{s[1]}")
def test_rsi16_mixed_plan_class_shapes(shape):
    """Encode lanes per codeword and LDS repair shapes set per class: the mixed fixtures unchanged."""
    from backends import PlanBackend
    R.check_rsi16_mixed(PlanBackend(graph_replays=2, shape=_rsi16_all(*shape)))


@pytest.mark.parametrize("shape", ["rs256", "rsi16"])
def test_plan_tuner_applies(shape, tmp_path, monkeypatch):
    """The plan tuner times every candidate shape of each class (checking they agree), caches the
    choice and applies it; the plan then still reproduces the mixed fixtures."""
    monkeypatch.setenv("FFRS_TUNE_CACHE", str(tmp_path / "tune.json"))
    from libffrs_amd import autotune
    from backends import PlanBackend

    def tuned(plan):
        res = autotune.apply_plan(plan, count=1024, reps=1)
        assert res
    be = PlanBackend(graph_replays=1, shape=tuned)
    if shape == "rs256":
        R.check_rs256_mixed(be)
    else:
        R.check_rsi16_mixed(be)
    assert (tmp_path / "tune.json").exists()
