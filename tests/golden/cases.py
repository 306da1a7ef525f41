"""Input builders for the golden fixtures, shared by make_golden.py and the tests.

Each builder regenerates its inputs deterministically from a seed (SplitMix64
stream, libffrs_amd/workloads.py); the fixture stores a SHA-256 of those inputs
so that drift in the generator is caught before any output is compared.
"""
import hashlib
import zlib

import numpy as np

from libffrs_amd import workloads as W


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), np.uint8).copy()


def row_crc(a):
    a = np.ascontiguousarray(a)
    a = a.reshape(a.shape[0], -1)
    return np.array([zlib.crc32(r.tobytes()) for r in a], np.uint32)


# ------------------------------------------------------------------ RS256

def rs256_plain_case(name):
    """(ecc, msgs[N,k], counts, stream) for the plain RS256 cases."""
    if name == "small":        # 4,096 x 247 bytes, ecc 8, 0-4 errors, seed 0
        msgs, counts, s = W.rs256_plain(4096, 8, 4, seed=0)
        return 8, msgs, counts, s
    if name == "scale_sub":    # the ecc-32 distribution at 8,192 codewords
        msgs, counts, s = W.rs256_plain(8192, 32, 16, seed=1, uncorrectable_frac=0.01)
        return 32, msgs, counts, s
    if name == "beyond_32":    # ecc 32 with 17-30 errors
        msgs, _, s = W.rs256_plain(2048, 32, 0, seed=2)
        counts = s.ints(17, 30, 2048)
        return 32, msgs, counts, s
    if name == "beyond_8":
        msgs, _, s = W.rs256_plain(2048, 8, 0, seed=3)
        counts = s.ints(5, 12, 2048)
        return 8, msgs, counts, s
    if name == "beyond_2":
        msgs, _, s = W.rs256_plain(2048, 2, 0, seed=4)
        counts = s.ints(2, 5, 2048)
        return 2, msgs, counts, s
    raise KeyError(name)


RS256_PLAIN_CASES = ("small", "scale_sub", "beyond_32", "beyond_8", "beyond_2")

# (block_len, ecc_len) shapes exercising every encoder word class and short blocks
RS256_SHAPES = ((255, 1), (255, 2), (255, 3), (10, 5), (255, 7), (255, 8), (200, 9), (255, 16), (40, 17),
                (255, 24), (255, 31), (255, 33), (255, 64), (100, 65), (255, 128), (255, 200), (255, 254))


def rs256_shape_case(block_len, ecc):
    """Buffer with full blocks plus a ragged tail, and 64 codewords for repair."""
    k = block_len - ecc
    s = W.Stream(1000 + block_len * 256 + ecc)
    nfull = 37
    tail = int(s.ints(1, k, 1)[0]) if k > 1 else 0
    buf = s.bytes(nfull * k + tail)
    rep_msgs = s.bytes(64 * k).reshape(64, k)
    counts = s.ints(0, ecc // 2 + 3, 64)
    return buf, rep_msgs, counts, s


def rs256_mixed_case(n=3000, seed=40):
    eccs, lens, counts, s = W.rs256_mixed(n, seed)
    msgs = [s.bytes(int(l)) for l in lens]
    return eccs, lens, counts, msgs, s


# ------------------------------------------------------------------ RSi16

def rsi16_plain_case(name):
    if name == "n256_sub":
        msgs, counts, s = W.rsi16_plain(4096, 256, 32, 16, seed=5)
        return (256, 32), msgs, counts, s
    if name == "n4096_sub":
        msgs, counts, s = W.rsi16_plain(64, 4096, 256, 128, seed=6)
        return (4096, 256), msgs, counts, s
    if name == "beyond_256":
        msgs, _, s = W.rsi16_plain(2048, 256, 32, 0, seed=7)
        counts = s.ints(17, 40, 2048)
        return (256, 32), msgs, counts, s
    if name == "beyond_64":
        msgs, _, s = W.rsi16_plain(2048, 64, 8, 0, seed=8)
        counts = s.ints(5, 12, 2048)
        return (64, 8), msgs, counts, s
    if name == "small_codes":
        msgs, counts, s = W.rsi16_plain(4096, 16, 4, 4, seed=9)
        return (16, 4), msgs, counts, s
    if name == "n1024_e128":
        msgs, counts, s = W.rsi16_plain(256, 1024, 128, 70, seed=10)
        return (1024, 128), msgs, counts, s
    if name == "n96_e32":      # block_len not a power of two (ntt_len 128)
        msgs, counts, s = W.rsi16_plain(1024, 96, 32, 18, seed=11)
        return (96, 32), msgs, counts, s
    raise KeyError(name)


RSI16_PLAIN_CASES = ("n256_sub", "n4096_sub", "beyond_256", "beyond_64", "small_codes", "n1024_e128", "n96_e32")


def rsi16_mixed_case(n=600, seed=41):
    ns, eccs, counts, s = W.rsi16_mixed(n, seed)
    msgs = [s.u16(int(nn - e)) for nn, e in zip(ns, eccs)]
    return ns, eccs, counts, msgs, s


def rbo(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


def rsi16_last_synd_zero(roots_ntt, n, ecc, count, seed=12):
    """Error patterns (positions, values) whose syndrome S_{ecc-1} is zero: the last
    error value is solved from the others (S_j(e) = sum e_m X_m^j, X_m = w^rbo(m))."""
    p = 65537
    N = len(roots_ntt)
    bits = N.bit_length() - 1
    s = W.Stream(seed)
    out = []
    while len(out) < count:
        c = int(s.ints(2, ecc // 2, 1)[0])
        pos = s.distinct_positions(np.array([c]), n)[0]
        vals = [int(v) for v in s.ints(1, 65536, c - 1)]
        X = [int(roots_ntt[rbo(int(m), bits)]) for m in pos]
        acc = sum(v * pow(x, ecc - 1, p) for v, x in zip(vals, X[:-1])) % p
        last = (-acc) * pow(pow(X[-1], ecc - 1, p), p - 2, p) % p
        if last == 0:
            continue
        out.append((pos.astype(np.int64), np.array(vals + [last], np.int64)))
    return out


def rsi16_radius_case(n_cw=65536, seed=16):
    """RSi16(256, 32) messages with 0-16 additive errors, searched for codewords the
    reference fails to repair although they carry at most ecc/2 errors."""
    return W.rsi16_plain(n_cw, 256, 32, 16, seed=seed)


def rsi16_known_case(n_cw=500, seed=15):
    """RSi16(256, 32) codewords with 0-28 additive errors, and the error positions to
    hand to repair: the corrupted positions plus up to 4 clean ones (at most 32)."""
    n, ecc = 256, 32
    msgs, counts, s = W.rsi16_plain(n_cw, n, ecc, 28, seed=seed)
    return msgs, counts, s


def rsi16_known_positions(pos, s, n=256):
    rows = pos.shape[0]
    extra = s.ints(0, 4, rows)
    allpos = np.full((rows, 32), -1, np.int16)
    for i in range(rows):
        p = [int(x) for x in pos[i] if x >= 0]
        cand = [int(x) for x in s.distinct_positions(np.array([8]), n)[0] if int(x) not in p]
        p = (p + cand[:int(extra[i])])[:32]
        allpos[i, :len(p)] = p
    return allpos
