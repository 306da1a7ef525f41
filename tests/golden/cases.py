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


# ------------------------------------------------------------------ shapes of the reference's own tests

# tests/test_lib_rs256.py:73-82 of the reference: block_len 239 and 255, ecc_len 1-16, 31-33, 63-65,
# 127-129, 253, 254 (where block_len > ecc_len)
RS256_REF_SHAPES = tuple((bl, e) for bl in (239, 255)
                         for e in list(range(1, 17)) + [31, 32, 33, 63, 64, 65, 127, 128, 129, 253, 254] if bl > e)


def rs256_reftest_case(block_len, ecc, count=24):
    """`count` messages of one reference test shape, with 0 .. ecc/2 + 2 errors each (some
    codewords lie beyond the decoding radius)."""
    k = block_len - ecc
    s = W.Stream(5000 + block_len * 256 + ecc)
    msgs = s.bytes(count * k).reshape(count, k)
    counts = s.ints(0, min(ecc // 2 + 2, block_len), count)
    return msgs, counts, s


# tests/test_lib_rsi16.py:399-455 of the reference (TestRSPower2, TestRSMult, TestRSLarge)
RSI16_REF_SHAPES = ((4, 2), (8, 2), (16, 4), (16, 8), (128, 2), (256, 32), (208, 8), (768, 8), (4096, 8),
                    (4096, 256), (2048, 128), (4096, 2048), (32768, 256), (65536, 256))


def rsi16_reftest_case(n, ecc):
    """Messages of one reference test shape (64 codewords for short blocks, down to 4 for
    n = 65536) with 0 .. ecc/2 + 1 additive errors each."""
    count = int(max(4, min(64, (1 << 18) // n)))
    s = W.Stream(7000 + n + ecc)
    msgs = s.u16(count * (n - ecc)).reshape(count, n - ecc)
    counts = s.ints(0, min(ecc // 2 + 1, n), count)
    return msgs, counts, s


# ------------------------------------------------------------------ SIMD-word parity corner

# Shapes with a mix constant -(w^(-rbo(k) j) / ecc) equal to 65536 (= -1) for some j (the
# mix of libffrs/rsi16v_impl.hpp:68-86 of the reference). There the SIMD words of the encode
# (galois.hpp:375-393: a residue that is 0 mod p and non-negative reduces to p, not 0, and
# p * 65536 wraps to 65536 in 32 bits) give other parity than the scalar words whenever the
# message residue at that j is 0 mod p, e.g. for the all-zero message.
RSI16_DIV_SHAPES = ((22, 2), (24, 4), (32, 8), (48, 8), (128, 8), (256, 128), (512, 128), (512, 256),
                    (4096, 1024))


def rsi16_mix_minus_one(n, ecc):
    """Indices j < ecc whose mix constant is 65536, for RSi16(n, ecc) with primitive 3."""
    p = 65537
    N = 1 << (n - 1).bit_length()
    bits = N.bit_length() - 1
    w = pow(3, 65536 // N, p)
    wm = pow(w, rbo(n - ecc, bits), p)
    wmi = pow(wm, p - 2, p)
    inv_e = pow(ecc, p - 2, p)
    return [j for j in range(ecc) if (-pow(wmi, j, p) * inv_e) % p == 65536]


def rsi16_message_residue(n, ecc, msg, j):
    """sum_m msg_m w^(j rbo(m)) mod p over the message symbols (the residue the mix scales)."""
    p = 65537
    N = 1 << (n - 1).bit_length()
    bits = N.bit_length() - 1
    w = pow(3, 65536 // N, p)
    acc = 0
    for m, v in enumerate(msg):
        if v:
            acc = (acc + int(v) * pow(w, (j * rbo(m, bits)) % N, p)) % p
    return acc


def rsi16_divergence_case(n, ecc):
    """Messages for RSi16(n, ecc): all zero, all 65535, messages whose residue at a
    mix-minus-one index j is 0 mod p (symbol 0, whose locator is 1, adjusted), and random
    ones."""
    k = n - ecc
    s = W.Stream(9000 + n * 16 + ecc)
    js = rsi16_mix_minus_one(n, ecc)
    rows = [np.zeros(k, np.uint16), np.full(k, 65535, np.uint16)]
    want = 6 if k <= 512 else 3
    made = 0
    while made < want:
        m = s.u16(k)
        j = js[made % len(js)]
        r = rsi16_message_residue(n, ecc, m, j)
        v = (int(m[0]) - r) % 65537
        if v == 65536:
            continue
        m[0] = v
        rows.append(m)
        made += 1
    rows += [s.u16(k), s.u16(k)]
    return np.stack(rows)


# ------------------------------------------------------------------ interleaved repair, SIMD words

# (block_len, ecc_len, interleave) of the coupled-group repair fixtures; groups of 16, 8 and 4
# columns, with a zero-filled last group where the interleave is not a multiple of the width
RSI16_ILSIMD_CASES = ((64, 8, 16), (64, 8, 21), (256, 32, 16), (256, 32, 7), (64, 16, 40), (1024, 128, 17),
                      (16, 4, 9), (4096, 256, 5))
RSI16_ILSIMD_WIDTHS = (16, 8, 4)
RSI16_ILSIMD_TRIALS = 4


def rsi16_ilsimd_messages(n, ecc, il, width, trial):
    """The interleaved message of one trial, and the stream that corrupts it."""
    s = W.Stream(11000 + 977 * n + 131 * ecc + 17 * il + 5 * width + trial)
    return s.u16((n - ecc) * il), s


def rsi16_ilsimd_corrupt(n, ecc, il, msg, par, s, trial):
    """Additive errors per column: 0 .. ecc/2 + 2 in even trials (some columns beyond the
    radius), 0 .. ecc/2 in odd ones."""
    k = n - ecc
    cws = np.concatenate([msg.reshape(k, il), par.reshape(ecc, il)], 0).T.copy()  # [il, n]
    counts = s.ints(0, ecc // 2 + (2 if trial % 2 == 0 else 0), il)
    W.rsi16_corrupt(cws, counts, s)
    return np.ascontiguousarray(cws[:, :k].T).reshape(-1), np.ascontiguousarray(cws[:, k:].T).reshape(-1)


RSI16_ILSIMD_KNOWN_TRIALS = 2


def rsi16_ilsimd_known(n, ecc, il, msg, par, s):
    """Known-position trials: 1 .. ecc positions, and in every column additive errors at a
    random subset of them. Returns (message, ecc, positions)."""
    k = n - ecc
    npos = int(s.ints(1, ecc, 1)[0])
    pos = np.sort(s.distinct_positions(np.array([npos]), n)[0]).astype(np.int64)
    cws = np.concatenate([msg.reshape(k, il), par.reshape(ecc, il)], 0).T.astype(np.int64)  # [il, n]
    for c in range(il):
        ne = int(s.ints(0, npos, 1)[0])
        where = pos[:ne]
        v = (cws[c, where] + s.ints(1, 65536, ne)) % 65537
        v[v == 65536] = 0
        cws[c, where] = v
    cws = cws.astype(np.uint16)
    return (np.ascontiguousarray(cws[:, :k].T).reshape(-1), np.ascontiguousarray(cws[:, k:].T).reshape(-1), pos)


# The NTT class and RSi16._sugiyama<W>: (block_len, ecc_len, primitive) and the vector counts
NTT_SHAPES = ((4, 2, 3), (8, 4, 3), (16, 8, 3), (32, 16, 3), (256, 32, 3), (1024, 64, 3), (16, 8, 5))
NTT_WIDTHS = (1, 4, 8, 16)
SUGIYAMA_SHAPES = ((16, 8), (64, 16), (256, 32), (1024, 64))


def ntt_inv_mods(E):
    """The n of the poly_inv trials: every n for E <= 16, else the powers of two, their
    neighbours and E."""
    if E <= 16:
        return list(range(0, E + 1))
    out = {0, 1, E}
    p = 2
    while p <= E:
        out.update({p - 1, p, min(p + 1, E)})
        p *= 2
    return sorted(out)


def poly_lengths(s, W, E, kind):
    """Per-vector polynomial lengths of a trial: 'half' (E/2, as the reference's own tests),
    'full', 'rand' (0..E) or 'pos' (1..E)."""
    if kind == "half":
        return np.full(W, E // 2, np.int64)
    if kind == "full":
        return np.full(W, E, np.int64)
    if kind == "rand":
        return s.ints(0, E, W)
    return s.ints(1, E, W)


def poly_vectors(s, W, E, lens):
    """W polynomials of E coefficients (u16, nonzero leading term), vector v of length lens[v]."""
    x = s.u16(W * E).reshape(W, E)
    for v in range(W):
        x[v, lens[v]:] = 0
        if lens[v] > 0 and x[v, lens[v] - 1] == 0:
            x[v, lens[v] - 1] = 1
    return x.reshape(-1)


# CIRC16: (inner_block_len, inner_ecc_len, outer_block_len, outer_ecc_len, interleave, SIMD width)
# -- the reference's own test shapes (tests/test_lib_circ16.py:253-300) with their flag variants
CIRC16_CASES = (
    (8, 2, 8, 2, 1, 16), (16, 4, 16, 4, 1, 16), (32, 16, 4, 2, 1, 16), (128, 8, 128, 8, 1, 16),
    (512, 4, 512, 32, 1, 16), (32768, 4, 16, 4, 1, 16), (16, 4, 32768, 4, 1, 16),
    (4, 2, 4, 2, 1, 8), (16, 4, 16, 4, 1, 8), (128, 8, 128, 8, 1, 8), (512, 4, 512, 32, 1, 8),
    (4, 2, 4, 2, 1, 4), (32, 16, 4, 2, 1, 4), (128, 8, 128, 8, 1, 4),
    (4, 2, 4, 2, 1, 1), (16, 4, 16, 4, 1, 1), (32, 16, 4, 2, 1, 1), (128, 8, 128, 8, 1, 1),
    (128, 2, 32, 2, 32, 16), (512, 4, 128, 8, 8, 16), (128, 8, 128, 8, 3, 16), (128, 8, 128, 8, 4, 16),
    (128, 8, 128, 8, 4, 8), (128, 8, 128, 8, 4, 4), (128, 8, 128, 8, 4, 1),
    (12, 2, 24, 2, 1, 16), (100, 2, 100, 2, 1, 16), (52, 2, 192, 64, 1, 16), (100, 2, 100, 2, 3, 16),
    (52, 2, 192, 64, 7, 16),
)


def circ16_flags(w):
    return {16: {}, 8: dict(simd_x16=False), 4: dict(simd_x16=False, simd_x8=False),
            1: dict(simd_x16=False, simd_x8=False, simd_x4=False)}[w]


def circ16_tag(case):
    return "_".join(str(x) for x in case)


class CircLayout:
    """Offsets of pycirc16.hpp:423-445 for one block."""

    def __init__(self, ibl, iecc, obl, oecc, il):
        self.ik, self.iE, self.ok, self.oE, self.on, self.I = ibl - iecc, iecc, obl - oecc, oecc, obl, il
        self.rI = self.ik * il
        self.msg_len = self.ok * self.rI
        self.rso_len = self.oE * self.rI
        self.ecc_len = self.rso_len + self.iE * self.on * il

    def message_offset(self, il, row, col):
        return self.rI * row + il * self.ik + col

    def rso_ecc_offset(self, il, row, col):
        return self.rI * row + il * self.ik + col

    def rsi_ecc_offset(self, il, row, col):
        return self.rso_len + (row * self.I + il) * self.iE + col


def circ16_xor(s, n):
    """n u16 masks with both bytes non-zero (the reference's tests XOR each byte with 1..255)."""
    lo = s.ints(1, 255, n)
    hi = s.ints(1, 255, n)
    return (lo | (hi << 8)).astype(np.uint16)


def circ16_corrupt_rows(lay, msg, ecc, rows, ils, s):
    """tests/test_lib_circ16.py:68-96: every symbol of the given outer rows (message or outer
    parity, and their inner parity) of the given interleave blocks XOR a mask."""
    for row in rows:
        for il in ils:
            if row < lay.ok:
                idx = [lay.message_offset(il, row, c) for c in range(lay.ik)]
                msg[idx] ^= circ16_xor(s, len(idx))
            else:
                idx = [lay.rso_ecc_offset(il, row - lay.ok, c) for c in range(lay.ik)]
                ecc[idx] ^= circ16_xor(s, len(idx))
            idx = [lay.rsi_ecc_offset(il, row, c) for c in range(lay.iE)]
            ecc[idx] ^= circ16_xor(s, len(idx))


def circ16_rows(s, lay, n):
    """n distinct outer rows."""
    return sorted(int(x) for x in s.distinct_positions([n], lay.on)[0])


def circ16_message(i, case):
    lay = CircLayout(*case[:5])
    return lay, W.Stream(1600 + i).u16(lay.msg_len)


def circ16_scenarios(i, case, lay, msg, ecc):
    """(name, message, ecc) repair inputs of case i: clean; the outer rows of block 0 corrupted
    (outer_ecc_len rows, as the reference's test_circ_repair); the same rows in every block; every
    inner parity of the message rows replaced and half as many rows corrupted
    (test_repair_fallback); more corrupted rows than outer_ecc_len (the repair without positions)."""
    s = W.Stream(1700 + i)
    out = [("none", msg.copy(), ecc.copy())]
    m, e = msg.copy(), ecc.copy()
    circ16_corrupt_rows(lay, m, e, circ16_rows(s, lay, lay.oE), [0], s)
    out.append(("rows", m, e))
    if lay.I > 1:
        m, e = msg.copy(), ecc.copy()
        circ16_corrupt_rows(lay, m, e, circ16_rows(s, lay, lay.oE), list(range(lay.I)), s)
        out.append(("rowsall", m, e))
    m, e = msg.copy(), ecc.copy()
    n_rsi = lay.ok * lay.I * lay.iE
    e[lay.rso_len:lay.rso_len + n_rsi] = s.u16(n_rsi)
    circ16_corrupt_rows(lay, m, e, circ16_rows(s, lay, max(1, lay.oE // 2)), [0], s)
    out.append(("fallback", m, e))
    m, e = msg.copy(), ecc.copy()
    circ16_corrupt_rows(lay, m, e, circ16_rows(s, lay, min(lay.on, lay.oE + 1 + lay.oE // 2)), [0], s)
    out.append(("many", m, e))
    return out


def circ16_find_inputs(i, case, lay, msg, ecc):
    """(interleave block, message) inputs of _find_outer_error_locations: the diagonal of block il
    corrupted (test_find_errors_diagonal) and one message row of it (test_find_errors_row)."""
    s = W.Stream(1800 + i)
    out = []
    for il in range(min(lay.I, 3)):
        m = msg.copy()
        for r in range(min(lay.ik, lay.ok)):
            m[lay.message_offset(il, r, r)] ^= circ16_xor(s, 1)[0]
        out.append((il, m))
        m = msg.copy()
        row = int(s.ints(0, lay.ok - 1, 1)[0])
        cols = s.distinct_positions([min(5, lay.ik)], lay.ik)[0]
        for c in cols:
            m[lay.message_offset(il, row, int(c))] ^= circ16_xor(s, 1)[0]
        out.append((il, m))
    return out


CIRC16_PLANT_CASES = ((8, 2, 8, 2, 1, 16), (16, 4, 16, 4, 1, 16), (16, 4, 16, 4, 1, 1), (4, 2, 4, 2, 1, 4),
                      (128, 8, 128, 8, 1, 16), (128, 8, 128, 8, 4, 8), (100, 2, 100, 2, 3, 16), (32, 16, 4, 2, 1, 1))


def circ16_planted_scenarios(i, kind, lay, msg, ecc):
    """Repair inputs of a planted block: clean, and outer_ecc_len rows of block 0 corrupted."""
    s = W.Stream(1950 + 2 * i + (kind == "inner"))
    m, e = msg.copy(), ecc.copy()
    circ16_corrupt_rows(lay, m, e, circ16_rows(s, lay, lay.oE), [0], s)
    return [("none", msg.copy(), ecc.copy()), ("rows", m, e)]


def circ16_planted_message(i, kind, lay, at, values):
    """A seeded message with `values` planted: outer column `at` (kind 'outer') or inner message
    row block `at` (kind 'inner')."""
    m = W.Stream(2000 + 2 * i + (kind == "inner")).u16(lay.msg_len)
    v = np.asarray(values, np.uint16)
    if kind == "outer":
        m[[lay.rI * r + at for r in range(lay.ok)]] = v
    else:
        m[at * lay.ik:(at + 1) * lay.ik] = v
    return m
