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


def test_rsi16_interleaved_repair_simd_words_raises(gpu):
    """With the SIMD words the reference repairs columns in coupled SIMD groups; not built."""
    import numpy as np
    ri = gpu.L.RSi16(64, 8, 3, simd_x4=True)
    assert ri.message_len == 56 * 3 and ri.ecc_len == 8 * 3 and ri.interleave == 3
    m = np.zeros(56 * 3, np.uint16)
    e = np.zeros(8 * 3, np.uint16)
    with pytest.raises(RuntimeError, match="interleave"):
        ri.repair(m.view(np.uint8), e.view(np.uint8))
