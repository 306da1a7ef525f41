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
