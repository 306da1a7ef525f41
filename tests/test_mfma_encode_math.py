"""The arithmetic of the int8 MFMA RSi16 encode (rsi16_encode_mfma_kernel), restated in numpy
and checked against the oracle on the CPU: the limb split of messages (2 bytes, as int8 after
XOR 0x80) and of G (three 7-bit limbs), the int32 limb products, the offset constant K_j and the
2^16 = -1 folds of the epilogue. G is the oracle's scalar-word encode of the unit messages, so
the check also pins that the encode is that linear map (for the scalar words always, for the
SIMD words on shapes whose mix has no 65536)."""
import numpy as np
import pytest

from libffrs_amd import workloads as W

P = 65537


def fold16(x):
    x = x.astype(np.int64)
    return ((x & 0xFFFF) - (x >> 16)).astype(np.int64)


def mfma_encode(G, msgs):
    """msgs [N, k] u16, G [k, E] canonical -> parity [N, E] u16, as the kernel computes it."""
    k, E = G.shape
    g = [(G & 127).astype(np.int64), ((G >> 7) & 127).astype(np.int64), (G >> 14).astype(np.int64)]
    assert all(((x >= 0) & (x <= 127)).all() for x in g)
    v = [(msgs & 0xFF).astype(np.int64) - 128, (msgs >> 8).astype(np.int64) - 128]  # int8 (byte ^ 0x80)
    C = [x.sum(0) for x in g]
    K = (128 * 257 * (C[0] + 128 * C[1] + 16384 * C[2])) % P
    out = np.zeros((msgs.shape[0], E), np.int64)
    z = []
    for b in range(3):
        p0, p1 = v[0] @ g[b], v[1] @ g[b]  # int32 accumulators of the MFMA
        assert np.abs(p0).max() < 2 ** 31 and np.abs(p1).max() < 2 ** 31
        z.append(fold16(fold16(p0) + 256 * fold16(p1)))
    V = z[0] + 128 * z[1] + 16384 * z[2] + K
    assert np.abs(V).max() < 2 ** 31
    f = fold16(V)
    f = np.where(f < 0, f + P, f)
    f = np.where(f >= P, f - P, f)
    out[:] = f & 0xFFFF
    return out.astype(np.uint16)


@pytest.mark.parametrize("n,ecc", [(256, 32), (64, 16), (1024, 128), (4096, 256), (208, 16)])
def test_mfma_encode_arithmetic(oracle_backend, n, ecc):
    k = n - ecc
    ri = oracle_backend.ri(n, ecc)
    # G row m = the encode of the unit message at m, exact (scalar words; & 0xFFFF would lose
    # 65536, so the canonical row is taken from the parity of 65536-free combinations below)
    eye = np.eye(k, dtype=np.uint16)
    Gt = ri.encode(eye, False).astype(np.int64)
    # a canonical entry of 65536 appears as 0 after & 0xFFFF: recover it from the encode of
    # 2 e_m, whose parity 2 g mod p is 65535 exactly when g = 65536
    G2 = ri.encode((2 * eye).astype(np.uint16), False).astype(np.int64)
    G = np.where((Gt == 0) & (G2 == 65535), 65536, Gt)
    msgs, _, _ = W.rsi16_plain(64, n, ecc, 0, seed=31)
    edge = np.stack([np.zeros(k, np.uint16), np.full(k, 65535, np.uint16)])
    msgs = np.concatenate([msgs, edge])
    want = ri.encode(msgs, False)
    got = mfma_encode(G, msgs)
    assert (got == want).all()
    # the SIMD words agree with the linear map on these shapes (no mix constant of 65536)
    assert (ri.encode(msgs, True) == want).all()


def test_mfma_encode_arithmetic_catches_a_wrong_offset(oracle_backend):
    """A dropped offset term gives other parity (the check above is not vacuous)."""
    n, ecc = 256, 32
    k = n - ecc
    ri = oracle_backend.ri(n, ecc)
    G = ri.encode(np.eye(k, dtype=np.uint16), False).astype(np.int64)
    msgs, _, _ = W.rsi16_plain(8, n, ecc, 0, seed=32)
    v = msgs.astype(np.int64)
    exact = (v @ G) % P & 0xFFFF
    assert (exact == ri.encode(msgs, False)).all()
    wrong = ((v - 128) @ G) % P & 0xFFFF
    assert (wrong != exact).any()
