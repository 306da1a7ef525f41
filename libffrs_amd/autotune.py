"""Kernel-shape tuning per (field, block_len, ecc_len), by timing on the device.

RSi16 encode and repair kernels run one codeword per group of g lanes (a power of two up
to min(64, ecc_len)); a wave holds 64 / g codewords. `tune_rsi16` times every legal g on
a synthetic batch, checks that every g gives the same parity, repaired symbols and
statuses, and returns the fastest for encode and for repair. Results are cached per
device architecture in a JSON file (FFRS_TUNE_CACHE, default
~/.cache/libffrs_amd/tune.json) and applied to a codec with `apply_rsi16`.

RS256 kernel shapes follow from ecc_len (slicing width, words per lane); the only choice,
pipeline or one codeword per lane for ecc_len <= 32, is measured by `tune_rs256`.
"""
import json
import os

import numpy as np

from . import capi
from . import workloads as W


def _cache_path():
    return os.environ.get("FFRS_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "libffrs_amd",
                                                          "tune.json"))


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache):
    path = _cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _key(ctx, *parts):
    arch = ctx.info()[0].split(" ")[0]
    return "/".join([arch] + [str(p) for p in parts])


def _time(ctx, fn, reps, before=None):
    best = float("inf")
    for _ in range(reps):
        if before:
            before()
        ctx.timer_begin()
        fn()
        best = min(best, ctx.timer_end())
    return best


def tune_rsi16(ctx, block_len, ecc_len, count=None, reps=3, use_cache=True):
    """Returns {"g_encode", "g_repair", "timings"} for RSi16(block_len, ecc_len) on ctx's
    device, from the cache when present."""
    key = _key(ctx, "rsi16", block_len, ecc_len)
    cache = _load() if use_cache else {}
    if key in cache:
        return cache[key]
    n, ecc = block_len, ecc_len
    k = n - ecc
    if count is None:
        count = max(256, min(65536, (16 << 20) // (2 * n)))
    msgs, counts, s = W.rsi16_plain(count, n, ecc, ecc // 2, seed=11)
    codec = capi.RSi16Codec(ctx, n, ecc)
    bufs = []

    def dev(a):
        b = ctx.alloc(a.nbytes)
        b.upload(np.ascontiguousarray(a))
        bufs.append(b)
        return b

    try:
        dm = dev(msgs)
        dp = ctx.alloc(count * ecc * 2)
        bufs.append(dp)
        codec.encode_device(dm, count, dp)
        par = np.empty((count, ecc), np.uint16)
        dp.download(par)
        cws = np.concatenate([msgs, par], 1)
        W.rsi16_corrupt(cws, counts, s)
        cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
        dcm, dcp = dev(cm), dev(cp)
        dst = ctx.alloc(count)
        bufs.append(dst)
        ref = None
        timings = {}
        g = 1
        while g <= min(64, ecc):
            try:
                codec.set_groups(g, g)
            except capi.FfrsError:
                g *= 2
                continue
            te = _time(ctx, lambda: codec.encode_device(dm, count, dp), reps)
            tr = _time(ctx, lambda: codec.repair_device(dcm, dcp, count, dst), reps,
                       before=lambda: (dcm.upload(cm), dcp.upload(cp)))
            out = [np.empty_like(par), np.empty_like(cm), np.empty_like(cp), np.empty(count, np.uint8)]
            for b, o in zip((dp, dcm, dcp, dst), out):
                b.download(o)
            if ref is None:
                ref = out
            elif not all((a == b).all() for a, b in zip(out, ref)):
                raise RuntimeError(f"RSi16({n}, {ecc}): results differ with {g} lanes per codeword")
            timings[g] = (te, tr)
            g *= 2
        best_e = min(timings, key=lambda x: timings[x][0])
        best_r = min(timings, key=lambda x: timings[x][1])
        res = {"g_encode": best_e, "g_repair": best_r, "count": count,
               "timings_ms": {str(x): [round(a, 4), round(b, 4)] for x, (a, b) in timings.items()}}
    finally:
        for b in bufs:
            b.free()
        codec.close()
    if use_cache:
        cache = _load()
        cache[key] = res
        _store(cache)
    return res


def apply_rsi16(codec, ctx=None):
    """Sets a capi.RSi16Codec's lanes per codeword from the cache (tuning on a miss)."""
    res = tune_rsi16(ctx or codec.ctx, codec.block_len, codec.ecc_len)
    codec.set_groups(res["g_encode"], res["g_repair"])
    return res


def tune_rs256(ctx, ecc_len, message_len=None, count=65536, reps=3, use_cache=True):
    """Times RS256 repair as pipeline (mode 1) and one codeword per lane (mode 2)."""
    k = message_len or 255 - ecc_len
    key = _key(ctx, "rs256", k + ecc_len, ecc_len, count)
    cache = _load() if use_cache else {}
    if key in cache:
        return cache[key]
    msgs, counts, s = W.rs256_plain(count, ecc_len, ecc_len // 2, seed=12)
    msgs = np.ascontiguousarray(msgs[:, :k])
    codec = capi.RS256Codec(ctx, ecc_len)
    dm = ctx.alloc(msgs.nbytes)
    dm.upload(msgs)
    dp = ctx.alloc(count * ecc_len)
    codec.encode_device(k, dm, count, dp)
    par = np.empty((count, ecc_len), np.uint8)
    dp.download(par)
    cws = np.concatenate([msgs, par], 1)
    W.rs256_corrupt(cws, counts, s)
    cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
    dst = ctx.alloc(count)
    timings = {}
    for mode in (1, 2):
        codec.set_repair_mode(mode)
        timings[mode] = _time(ctx, lambda: codec.repair_device(k, dm, dp, count, dst), reps,
                              before=lambda: (dm.upload(cm), dp.upload(cp)))
    for b in (dm, dp, dst):
        b.free()
    codec.close()
    res = {"repair_mode": min(timings, key=timings.get), "timings_ms": {str(m): round(t, 4) for m, t in timings.items()}}
    if use_cache:
        cache = _load()
        cache[key] = res
        _store(cache)
    return res


# ---------------------------------------------------------------- fused plans

def _flat_rows(rows, width):
    """Pads byte/u16 rows of different lengths into one packed buffer; returns (buffer, offsets)."""
    lens = np.array([len(r) for r in rows], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64)
    buf = np.concatenate(rows) if rows else np.zeros(0, width)
    return np.ascontiguousarray(buf, width), off


def _bucket(count):
    """Class sizes are tuned at the next power of two (1,024 to 65,536)."""
    b = 1024
    while b < count and b < 65536:
        b *= 2
    return b


def tune_rs256_plan_class(ctx, ecc_len, count=4096, reps=3, use_cache=True):
    """Times the repair of `count` codewords of ecc_len (message lengths uniform in
    [1, 255 - ecc_len], errors up to ecc_len / 2) as a one-class plan, one codeword per lane
    (mode 1) and per wave (mode 2); checks that both give the same bytes and statuses and
    returns {"mode", "timings_ms"} (cached per device architecture, ecc_len and count: which
    kernel wins depends on how many codewords share the device)."""
    key = _key(ctx, "rs256plan", ecc_len, count)
    cache = _load() if use_cache else {}
    if key in cache:
        return cache[key]
    s = W.Stream(40 + ecc_len)
    u = s.u64(count) >> np.uint64(32)
    lens = 1 + ((u * np.uint64(255 - ecc_len)) >> np.uint64(32)).astype(np.int64)
    msgs = [s.bytes(int(k)) for k in lens]
    mbuf, moff = _flat_rows(msgs, np.uint8)
    eoff = (np.arange(count, dtype=np.uint64) * np.uint64(ecc_len))
    plan = capi.RS256Plan(ctx, np.full(count, ecc_len), lens, moff, eoff)
    dm, dp, dst = ctx.alloc(max(1, mbuf.nbytes)), ctx.alloc(count * ecc_len), ctx.alloc(count)
    try:
        dm.upload(mbuf)
        plan.encode(dm, dp)
        par = np.empty(count * ecc_len, np.uint8)
        dp.download(par)
        counts = ((s.u64(count) >> np.uint64(32)) * np.uint64(ecc_len // 2 + 1) >> np.uint64(32)).astype(np.int64)
        cm, cp = mbuf.copy(), par.copy()
        for i in range(count):
            k = int(lens[i])
            cw = np.concatenate([cm[int(moff[i]):int(moff[i]) + k], cp[i * ecc_len:(i + 1) * ecc_len]])
            W.rs256_corrupt(cw[None, :], counts[i:i + 1], s)
            cm[int(moff[i]):int(moff[i]) + k], cp[i * ecc_len:(i + 1) * ecc_len] = cw[:k], cw[k:]
        timings, ref = {}, None
        for mode in (1, 2):
            plan.set_class_mode(ecc_len, mode)
            timings[mode] = _time(ctx, lambda: plan.repair(dm, dp, dst), reps,
                                  before=lambda: (dm.upload(cm), dp.upload(cp)))
            out = [np.empty_like(cm), np.empty_like(cp), np.empty(count, np.uint8)]
            for b, o in zip((dm, dp, dst), out):
                b.download(o)
            if ref is None:
                ref = out
            elif not all((a == b).all() for a, b in zip(out, ref)):
                raise RuntimeError(f"RS256 plan, ecc {ecc_len}: results differ between repair kernels")
    finally:
        for b in (dm, dp, dst):
            b.free()
        plan.close()
    res = {"mode": min(timings, key=timings.get), "timings_ms": {str(m): round(t, 4) for m, t in timings.items()}}
    if use_cache:
        cache = _load()
        cache[key] = res
        _store(cache)
    return res


def tune_rsi16_plan_class(ctx, block_len, ecc_len, count=None, reps=3, use_cache=True):
    """Times a one-class RSi16 plan of RSi16(block_len, ecc_len): encode with every legal lanes
    per codeword (ecc_len <= 32), repair with the register code and every LDS shape; checks that
    all give the same results; returns {"l_encode", "g_repair", "timings_ms"} (cached)."""
    key = _key(ctx, "rsi16plan", block_len, ecc_len, count)
    cache = _load() if use_cache else {}
    if key in cache:
        return cache[key]
    n, ecc = block_len, ecc_len
    k = n - ecc
    if count is None:
        count = max(256, min(16384, (8 << 20) // (2 * n)))
    msgs, counts, s = W.rsi16_plain(count, n, ecc, ecc // 2, seed=13)
    moff = np.arange(count, dtype=np.uint64) * np.uint64(k)
    eoff = np.arange(count, dtype=np.uint64) * np.uint64(ecc)
    plan = capi.RSi16Plan(ctx, np.full(count, n), np.full(count, ecc), moff, eoff)
    dm, dp, dst = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc * 2), ctx.alloc(count)
    try:
        dm.upload(np.ascontiguousarray(msgs))
        l_opts = [1, 2, 4, 8, 16, 32, 64] if ecc <= 32 else [0]
        timings_e, ref_e = {}, None
        for L in l_opts:
            plan.set_class_groups(n, ecc, L, 0)
            timings_e[L] = _time(ctx, lambda: plan.encode(dm, dp), reps)
            par = np.empty((count, ecc), np.uint16)
            dp.download(par)
            if ref_e is None:
                ref_e = par
            elif not (par == ref_e).all():
                raise RuntimeError(f"RSi16({n}, {ecc}) plan: parity differs with {L} encode lanes")
        cws = np.concatenate([msgs, ref_e], 1)
        W.rsi16_corrupt(cws, counts, s)
        cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
        g_opts = [0] + [g for g in (1, 2, 4, 8, 16, 32, 64) if g <= min(64, ecc) and g != min(64, max(1, ecc // 2))
                        and (64 // g) * 9 * ecc * 4 <= 160 * 1024]
        timings_r, ref_r = {}, None
        best_l = min(timings_e, key=timings_e.get)
        for g in g_opts:
            plan.set_class_groups(n, ecc, best_l, g)
            timings_r[g] = _time(ctx, lambda: plan.repair(dm, dp, dst), reps,
                                 before=lambda: (dm.upload(cm), dp.upload(cp)))
            out = [np.empty_like(cm), np.empty_like(cp), np.empty(count, np.uint8)]
            for b, o in zip((dm, dp, dst), out):
                b.download(o)
            if ref_r is None:
                ref_r = out
            elif not all((a == b).all() for a, b in zip(out, ref_r)):
                raise RuntimeError(f"RSi16({n}, {ecc}) plan: repair differs with {g} lanes per codeword")
    finally:
        for b in (dm, dp, dst):
            b.free()
        plan.close()
    res = {"l_encode": best_l, "g_repair": min(timings_r, key=timings_r.get),
           "timings_ms": {"encode": {str(x): round(t, 4) for x, t in timings_e.items()},
                          "repair": {str(x): round(t, 4) for x, t in timings_r.items()}}}
    if use_cache:
        cache = _load()
        cache[key] = res
        _store(cache)
    return res


def apply_plan(plan, ctx=None, **kw):
    """Sets every class of a capi.RS256Plan or capi.RSi16Plan to the tuned shape (from the cache,
    tuning on a miss at the class's own size in the plan, rounded to a power of two; keyword
    arguments go to the tuner). Returns {class: result}."""
    ctx = ctx or plan.ctx
    out = {}
    sizes = plan.class_sizes()
    if isinstance(plan, capi.RS256Plan):
        for e in plan.classes():
            r = tune_rs256_plan_class(ctx, e, **{"count": _bucket(sizes[e]), **kw})
            plan.set_class_mode(e, r["mode"])
            out[e] = r
    else:
        for n, e in plan.classes():
            r = tune_rsi16_plan_class(ctx, n, e, **{"count": _bucket(sizes[(n, e)]), **kw})
            plan.set_class_groups(n, e, r["l_encode"], r["g_repair"])
            out[(n, e)] = r
    return out
