"""Deterministic synthetic workloads for the RS256 / RSi16 paths.

All randomness comes from a counter-based SplitMix64 stream, so every input is
regenerated bit-identically on any host from (workload, seed) alone; nothing
depends on numpy's own generator algorithms.

Workloads (sizes as used by the benchmark and the tests):
  * rs256_plain_small  RS256 ecc 8, 4,096 messages of 247 bytes, 0-4 errors (XOR)
  * rs256_plain_scale  RS256 ecc 32, 1,048,576 messages of 223 bytes, 0-16 errors,
                       plus 1% of codewords with 17-20 errors
  * rsi16_plain        RSi16 n 256 ecc 32 (262,144 codewords) and n 4096 ecc 256
                       (16,384 codewords), additive errors mod 65537
  * mixed batch        RS256 ecc in {2,4,8,16,32,64} with message length uniform in
                       [1, 255-ecc]; RSi16 n in {64,256,1024}, ecc in {8,16,32}
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Stream:
    """SplitMix64 counter stream: value i is mix(seed + (i + 1) * golden)."""

    def __init__(self, seed):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.ctr = 0

    def u64(self, n):
        n = int(n)
        idx = np.arange(self.ctr + 1, self.ctr + 1 + n, dtype=np.uint64)
        self.ctr += n
        with np.errstate(over="ignore"):
            return _mix(self.seed + idx * _GOLDEN)

    def bytes(self, n):
        n = int(n)
        w = self.u64((n + 7) // 8)
        return w.view(np.uint8)[:n].copy()

    def ints(self, lo, hi, n):
        """n integers uniform in [lo, hi] (inclusive), by 32x32 multiply-high."""
        span = np.uint64(int(hi) - int(lo) + 1)
        r = self.u64(n) >> np.uint64(32)
        return (np.int64(lo) + ((r * span) >> np.uint64(32)).astype(np.int64))

    def u16(self, n):
        n = int(n)
        w = self.u64((n + 3) // 4)
        return w.view(np.uint16)[:n].copy()

    def distinct_positions(self, counts, n_positions):
        """For each row i, counts[i] distinct positions in [0, n_positions).

        Returns an int32 array [rows, max(counts)] padded with -1. Positions are drawn
        uniformly with replacement; every later duplicate within a row is redrawn until
        the row holds distinct values (sampling without replacement).
        """
        counts = np.asarray(counts, np.int64)
        rows = counts.size
        cmax = int(counts.max()) if rows else 0
        if cmax == 0:
            return np.full((rows, 0), -1, np.int32)
        if cmax > n_positions:
            raise ValueError("more errors than positions")
        valid = np.arange(cmax)[None, :] < counts[:, None]
        pos = self.ints(0, n_positions - 1, rows * cmax).reshape(rows, cmax)
        pos = np.where(valid, pos, -1 - np.arange(cmax)[None, :])  # padding never collides
        todo = np.arange(rows)
        while todo.size:
            sub = pos[todo]
            order = np.argsort(sub, axis=1, kind="stable")
            srt = np.take_along_axis(sub, order, axis=1)
            dup_sorted = np.zeros_like(srt, dtype=bool)
            dup_sorted[:, 1:] = srt[:, 1:] == srt[:, :-1]
            dup = np.zeros_like(dup_sorted)
            np.put_along_axis(dup, order, dup_sorted, axis=1)
            r, c = np.nonzero(dup)
            if r.size == 0:
                break
            sub[r, c] = self.ints(0, n_positions - 1, r.size)
            pos[todo] = sub
            todo = todo[np.unique(r)]
        return np.where(valid, pos, -1).astype(np.int32)


def rs256_corrupt(cws, counts, stream):
    """XOR counts[i] distinct bytes of codeword i with uniform non-zero bytes (in place)."""
    rows, n = cws.shape
    pos = stream.distinct_positions(counts, n)
    vals = stream.ints(1, 255, pos.size).reshape(pos.shape).astype(np.uint8)
    r, c = np.nonzero(pos >= 0)
    cws[r, pos[r, c]] ^= vals[r, c]
    return pos


def rsi16_corrupt(cws, counts, stream):
    """Add uniform non-zero offsets mod 65537 at counts[i] distinct symbols of codeword i,
    resampling an offset while the result would be 65536 (not representable in u16)."""
    rows, n = cws.shape
    pos = stream.distinct_positions(counts, n)
    r, c = np.nonzero(pos >= 0)
    p = pos[r, c]
    v = cws[r, p].astype(np.int64)
    o = stream.ints(1, 65536, v.size)
    nv = (v + o) % 65537
    bad = np.nonzero(nv == 65536)[0]
    while bad.size:
        o2 = stream.ints(1, 65536, bad.size)
        nv[bad] = (v[bad] + o2) % 65537
        bad = bad[nv[bad] == 65536]
    cws[r, p] = nv.astype(np.uint16)
    return pos


def rs256_plain(n_codewords, ecc_len, max_err, seed, uncorrectable_frac=0.0, uncorrectable_range=(17, 20)):
    """Messages [N, 255-ecc] and per-codeword error counts for an RS256 plain batch."""
    k = 255 - ecc_len
    s = Stream(seed)
    msgs = s.bytes(n_codewords * k).reshape(n_codewords, k)
    counts = s.ints(0, max_err, n_codewords)
    if uncorrectable_frac > 0:
        n_unc = int(round(n_codewords * uncorrectable_frac))
        # a fixed stride spreads the uncorrectable codewords over the batch
        idx = (np.arange(n_unc, dtype=np.int64) * n_codewords) // max(n_unc, 1)
        counts[idx] = s.ints(uncorrectable_range[0], uncorrectable_range[1], n_unc)
    return msgs, counts, s


def rsi16_plain(n_codewords, block_len, ecc_len, max_err, seed):
    k = block_len - ecc_len
    s = Stream(seed)
    msgs = s.u16(n_codewords * k).reshape(n_codewords, k)
    counts = s.ints(0, max_err, n_codewords)
    return msgs, counts, s


RS256_MIXED_ECC = (2, 4, 8, 16, 32, 64)
RSI16_MIXED_N = (64, 256, 1024)
RSI16_MIXED_ECC = (8, 16, 32)


def rs256_mixed(n_items, seed):
    """Per item: ecc in RS256_MIXED_ECC, message length in [1, 255-ecc], errors in
    [0, ecc/2]. Returns (eccs, lens, counts, stream)."""
    s = Stream(seed)
    eccs = np.array(RS256_MIXED_ECC)[s.ints(0, len(RS256_MIXED_ECC) - 1, n_items)]
    u = s.u64(n_items) >> np.uint64(32)
    lens = 1 + ((u * (255 - eccs).astype(np.uint64)) >> np.uint64(32)).astype(np.int64)
    u = s.u64(n_items) >> np.uint64(32)
    counts = ((u * (eccs // 2 + 1).astype(np.uint64)) >> np.uint64(32)).astype(np.int64)
    return eccs.astype(np.int64), lens, counts, s


def rsi16_mixed(n_items, seed):
    s = Stream(seed)
    ns = np.array(RSI16_MIXED_N)[s.ints(0, len(RSI16_MIXED_N) - 1, n_items)]
    eccs = np.array(RSI16_MIXED_ECC)[s.ints(0, len(RSI16_MIXED_ECC) - 1, n_items)]
    u = s.u64(n_items) >> np.uint64(32)
    counts = ((u * (eccs // 2 + 1).astype(np.uint64)) >> np.uint64(32)).astype(np.int64)
    return ns.astype(np.int64), eccs.astype(np.int64), counts, s


SERVE_MIN_BYTES, SERVE_MAX_BYTES = 4096, 64 << 20


def serving_requests(n_requests, rate, seed, min_bytes=SERVE_MIN_BYTES, max_bytes=SERVE_MAX_BYTES,
                     encode_frac=0.7):
    """Poisson arrivals at `rate` per second: (arrival s, kind, size bytes) per request, sizes
    log-uniform in [min_bytes, max_bytes], `encode_frac` of them encode and the rest
    verify+repair."""
    s = Stream(seed)
    u = (s.u64(n_requests) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    gaps = -np.log1p(-u) / rate
    t = np.cumsum(gaps) - gaps[0]
    v = (s.u64(n_requests) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    sizes = np.floor(min_bytes * (max_bytes / min_bytes) ** v).astype(np.int64)
    w = (s.u64(n_requests) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    kinds = np.where(w < encode_frac, "encode", "repair")
    return t, kinds, sizes, s


def serving_corrupt(msgs, parity, stream, max_frac=0.05):
    """Corrupts a request's codewords in place: a fraction uniform in [0, max_frac] of all
    symbols (message and parity, codeword-major positions), half as single-symbol errors
    and half as bursts of 8-64 contiguous symbols, each an additive non-zero offset mod
    65537 (resampled while the result would be 65536). Returns the number of corrupted
    symbol writes."""
    blocks, k = msgs.shape
    ecc = parity.shape[1]
    n = k + ecc
    total = blocks * n
    frac = float((stream.u64(1)[0] >> np.uint64(11))) / float(1 << 53) * max_frac
    count = int(round(frac * total))
    singles = count // 2
    pos = [stream.ints(0, total - 1, singles)] if singles else []
    burst_budget = count - singles
    while burst_budget > 0:
        ln = int(stream.ints(8, 64, 1)[0])
        ln = min(ln, burst_budget, total)
        start = int(stream.ints(0, total - ln, 1)[0])
        pos.append(np.arange(start, start + ln, dtype=np.int64))
        burst_budget -= ln
    if not pos:
        return 0
    pos = np.concatenate(pos)
    b, j = pos // n, pos % n
    flat = np.concatenate([msgs, parity], 1)
    v = flat[b, j].astype(np.int64)
    o = stream.ints(1, 65536, v.size)
    nv = (v + o) % 65537
    bad = np.nonzero(nv == 65536)[0]
    while bad.size:
        nv[bad] = (v[bad] + stream.ints(1, 65536, bad.size)) % 65537
        bad = bad[nv[bad] == 65536]
    flat[b, j] = nv.astype(np.uint16)
    msgs[:] = flat[:, :k]
    parity[:] = flat[:, k:]
    return int(pos.size)
