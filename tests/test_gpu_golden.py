"""The HIP path against the golden fixtures written from the reference's own code."""
import pytest

import cases as C
import replay as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from backends import GpuBackend
    return GpuBackend()


# RS256 repair: automatic kernel choice, the three-kernel pipeline, one kernel per lane
@pytest.fixture(scope="module", params=[0, 1, 2], ids=["auto", "pipeline", "lane"])
def gpu_rs(request):
    from backends import GpuBackend
    return GpuBackend(repair_mode=request.param)


@pytest.mark.parametrize("name", C.RS256_PLAIN_CASES)
def test_rs256_plain(gpu_rs, name):
    R.check_rs256_plain(gpu_rs, name)


@pytest.mark.parametrize("shape", C.RS256_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rs256_shapes(gpu_rs, shape):
    R.check_rs256_shape(gpu_rs, *shape)


def test_rs256_mixed(gpu_rs):
    R.check_rs256_mixed(gpu_rs)


@pytest.mark.parametrize("name", C.RSI16_PLAIN_CASES)
def test_rsi16_plain(gpu, name):
    R.check_rsi16_plain(gpu, name)


def test_rsi16_mixed(gpu):
    R.check_rsi16_mixed(gpu)


def test_rsi16_edges(gpu):
    R.check_rsi16_edges(gpu)


def test_rsi16_known(gpu):
    R.check_rsi16_known(gpu)


def test_rsi16_interleaved(gpu):
    R.check_rsi16_interleaved(gpu)


def test_rsi16_interleaved_repair(gpu):
    R.check_rsi16_interleaved_repair(gpu)


@pytest.mark.parametrize("shape", C.RS256_REF_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rs256_reference_test_shapes(gpu_rs, shape):
    R.check_rs256_reftest(gpu_rs, *shape)


@pytest.mark.parametrize("shape", C.RSI16_REF_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rsi16_reference_test_shapes(gpu, shape):
    R.check_rsi16_reftest(gpu, *shape)


@pytest.mark.parametrize("shape", C.RSI16_DIV_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rsi16_simd_word_divergence(gpu, shape):
    R.check_rsi16_divergence(gpu, *shape)


@pytest.mark.parametrize("width", C.RSI16_ILSIMD_WIDTHS)
@pytest.mark.parametrize("case", C.RSI16_ILSIMD_CASES, ids=lambda c: f"{c[0]}_{c[1]}_{c[2]}")
def test_rsi16_interleaved_repair_simd_words(gpu, case, width):
    """Coupled groups of 16, 8 or 4 columns, as the reference's vector instantiations repair them."""
    R.check_rsi16_ilsimd(gpu, *case, width)
