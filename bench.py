"""Benchmark: Reed-Solomon encode + repair GB/s on one MI355X (plain GPU path).

Default workload rs256_plain_scale: RS256 ecc_len 32, n 255 (223 message bytes), 1,048,576
codewords (233,832,448 message bytes); every codeword is corrupted at 0-16 distinct bytes,
1% at 17-20 (uncorrectable path). Other workloads (--workload): rs256_plain_small (ecc 8,
4,096 codewords), rs256_ecc8_scale (ecc 8, 4,194,304 codewords), rsi16_plain (RSi16(256, 32),
262,144 codewords) and rsi16_large (RSi16(4096, 256), 16,384 codewords). A step is one encode of all messages plus one repair
of all corrupted codewords, with inputs resident in HBM. The corrupted input is restored
from a pristine device copy before each step, outside the timed bracket.

    python bench.py [--workload W] [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

Prints one JSON line; `value` is message GB/s of the whole step, `roofline` reports the
dominant launch (its kernels timed with HIP events on the codec's stream), `cpu_baseline`
times the reference's own CPU build (or the oracle port where it is absent) on a bounded
sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RS encode/decode GB/s on 1 MI355X, fused-batched vs plain; p99 request latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ensure_built():
    import importlib.util
    spec = importlib.util.spec_from_file_location("_ffrs_build", os.path.join(ROOT, "libffrs_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    if not b.is_built():
        log("building libffrs_amd ...")
        b.build()


def load_reference_cpu():
    """The reference's own extension compiled from its sources (oracle/_ref), or None."""
    import importlib.util
    import sysconfig
    path = os.path.join(ROOT, "oracle", "_ref", "libffrs" + sysconfig.get_config_var("EXT_SUFFIX"))
    if not os.path.exists(path):
        return None
    try:
        spec = importlib.util.spec_from_file_location("libffrs", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        return mod
    except Exception as e:  # a host without the ISA it was built for, etc.
        log("reference CPU build not loadable:", e)
        return None


WORKLOADS = {
    # RS256 ecc 32 over 1,048,576 codewords of 223 bytes; 0-16 errors, 1% with 17-20
    "rs256_plain_scale": dict(field="rs256", N=1048576, n=255, ecc=32, max_err=16, unc=0.01, seed=2,
                              errors="0-16 per codeword, 1% with 17-20", cpu_repair=262144),
    # RS256 ecc 8 over 4,096 codewords of 247 bytes; 0-4 errors
    "rs256_plain_small": dict(field="rs256", N=4096, n=255, ecc=8, max_err=4, unc=0.0, seed=0,
                              errors="0-4 per codeword", cpu_repair=4096),
    # the same shape at scale: 4,194,304 codewords (1,036,017,664 message bytes)
    "rs256_ecc8_scale": dict(field="rs256", N=4194304, n=255, ecc=8, max_err=4, unc=0.0, seed=0,
                             errors="0-4 per codeword", cpu_repair=262144),
    # RSi16(256, 32) over 262,144 codewords; 0-16 additive errors
    "rsi16_plain": dict(field="rsi16", N=262144, n=256, ecc=32, max_err=16, seed=3,
                        errors="0-16 additive per codeword", cpu_repair=32768),
    # RSi16(4096, 256) over 16,384 codewords; 0-128 additive errors
    "rsi16_large": dict(field="rsi16", N=16384, n=4096, ecc=256, max_err=128, seed=3,
                        errors="0-128 additive per codeword", cpu_repair=256),
}


# Worker state of the all-cores CPU baseline: set in the parent before the pool forks, read by
# the workers through copy-on-write memory (nothing is pickled).
_CPU = {}


def _cpu_encode_shard(i):
    """Worker i: the reference's encode of its contiguous shard of the messages, one call."""
    ref, wl, msgs, nw = _CPU["ref"], _CPU["wl"], _CPU["msgs"], _CPU["workers"]
    n, ecc = wl["n"], wl["ecc"]
    a, b = len(msgs) * i // nw, len(msgs) * (i + 1) // nw
    codec = ref.RS256(block_len=n, ecc_len=ecc) if wl["field"] == "rs256" else ref.RSi16(n, ecc)
    buf = msgs[a:b].tobytes()
    t = time.perf_counter()
    codec.encode(buf)
    return time.perf_counter() - t


def _cpu_repair_shard(i):
    """Worker i: one repair call per corrupted codeword of its shard (the reference's API repairs
    one codeword per call)."""
    ref, wl, cws, nw = _CPU["ref"], _CPU["wl"], _CPU["cws"], _CPU["workers"]
    n, ecc = wl["n"], wl["ecc"]
    k = n - ecc
    a, b = len(cws) * i // nw, len(cws) * (i + 1) // nw
    codec = ref.RS256(block_len=n, ecc_len=ecc) if wl["field"] == "rs256" else ref.RSi16(n, ecc)
    mb = [bytearray(r[:k].tobytes()) for r in cws[a:b]]
    pb = [bytearray(r[k:].tobytes()) for r in cws[a:b]]
    t = time.perf_counter()
    for m, q in zip(mb, pb):
        try:
            codec.repair(m, q)
        except RuntimeError:
            pass
    return time.perf_counter() - t


def _cpu_repair_il16_shard(i):
    """Worker i: RSi16 repair with interleave 16, the reference's best CPU repair: 16 codewords
    per call, their columns repaired together by the 16-lane SIMD build."""
    ref, wl, cws, nw = _CPU["ref"], _CPU["wl"], _CPU["cws"], _CPU["workers"]
    n, ecc = wl["n"], wl["ecc"]
    k = n - ecc
    groups = len(cws) // 16
    a, b = groups * i // nw, groups * (i + 1) // nw
    codec = ref.RSi16(n, ecc, 16)
    bufs = []
    for g in range(a, b):
        blk = cws[16 * g:16 * g + 16]  # [16, n]: column c = codeword c
        bufs.append((bytearray(np.ascontiguousarray(blk[:, :k].T).tobytes()),
                     bytearray(np.ascontiguousarray(blk[:, k:].T).tobytes())))
    t = time.perf_counter()
    for m, q in bufs:
        try:
            codec.repair(m, q)
        except RuntimeError:
            pass
    return time.perf_counter() - t


def cpu_baseline(wl, msgs, sample):
    """The reference's own CPU build (oracle/_ref; the oracle port where that build is absent)
    on the host cores, before anything touches the GPU (the pool's workers are forked): encode
    of the message buffer and repair of `sample` corrupted codewords, one call each, first on
    one core, then split over min(16, host CPUs) worker processes (the reference is
    single-threaded; separate processes on disjoint shards are its all-cores form). For RSi16
    the repair is also timed through interleave 16 (16 codewords per call, the 16-lane SIMD
    build). Returns message GB/s of one encode + repair pass over the workload."""
    from libffrs_amd import workloads as W

    n, ecc = wl["n"], wl["ecc"]
    k = n - ecc
    elem = 1 if wl["field"] == "rs256" else 2
    ref = load_reference_cpu()
    kind = "reference" if ref is not None else "port"
    N = msgs.shape[0]
    # corrupted codewords of the sample, encoded by the same build and corrupted as the workload
    t0 = time.perf_counter()
    if ref is not None:
        codec = ref.RS256(block_len=n, ecc_len=ecc) if elem == 1 else ref.RSi16(n, ecc)
        par = np.frombuffer(codec.encode(msgs[:sample].tobytes()), msgs.dtype).reshape(sample, ecc)
    else:
        from oracle import oracle as O
        par = (O.RS256(ecc).encode_batch(msgs[:sample]) if elem == 1 else O.RSi16(n, ecc).encode(msgs[:sample], True))
    cws = np.concatenate([msgs[:sample], par], 1)
    if elem == 1:
        _, counts, s = W.rs256_plain(sample, ecc, wl["max_err"], seed=wl["seed"] + 100,
                                     uncorrectable_frac=wl.get("unc", 0.0))
        W.rs256_corrupt(cws, counts, s)
    else:
        _, counts, s = W.rsi16_plain(sample, n, ecc, wl["max_err"], seed=wl["seed"] + 100)
        W.rsi16_corrupt(cws, counts, s)
    log(f"CPU baseline sample of {sample} codewords prepared in {time.perf_counter() - t0:.1f} s")
    if ref is None:
        from oracle import oracle as O
        t = time.perf_counter()
        if elem == 1:
            O.RS256(ecc).encode_batch(msgs, threads=1)
        else:
            O.RSi16(n, ecc).encode(msgs, True, threads=1)
        te = time.perf_counter() - t
        m, q = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
        t = time.perf_counter()
        (O.RS256(ecc).decode_batch(m, q, threads=1) if elem == 1 else O.RSi16(n, ecc).repair_batch(m, q, threads=1))
        tr = time.perf_counter() - t
        per_byte = te / (N * k * elem) + tr / (sample * k * elem)
        return {"value": round(1.0 / per_byte / 1e9, 6), "unit": "GB/s", "cores": 1, "kind": kind,
                "sample": f"oracle port: encode of {N} messages + repair of {sample} codewords, one thread"}
    import multiprocessing as mp
    _CPU.update(ref=ref, wl=wl, msgs=msgs, cws=cws, workers=1)
    te1 = _cpu_encode_shard(0)
    tr1 = _cpu_repair_shard(0)
    workers = max(1, min(16, os.cpu_count() or 1))
    _CPU["workers"] = workers
    with mp.get_context("fork").Pool(workers) as pool:
        ten = max(pool.map(_cpu_encode_shard, range(workers)))
        trn = max(pool.map(_cpu_repair_shard, range(workers)))
        il16 = None
        if elem == 2 and sample >= 16 * workers:
            il16 = max(pool.map(_cpu_repair_il16_shard, range(workers)))
    _CPU.clear()
    enc_b, rep_b = N * k * elem, sample * k * elem

    def gb(b, t):
        return round(b / t / 1e9, 6)

    per_byte_1 = te1 / enc_b + tr1 / rep_b
    per_byte_n = ten / enc_b + trn / rep_b
    res = {
        "value": round(1.0 / per_byte_n / 1e9, 6),
        "unit": "GB/s",
        "cores": workers,
        "kind": kind,
        "sample": (f"encode of all {N} messages (one call per worker shard) + repair of {sample} corrupted "
                   f"codewords, one call each; {workers} worker processes on disjoint shards "
                   f"({os.cpu_count()} host CPUs), and one core alone"),
        "encode_gbps": gb(enc_b, ten),
        "repair_gbps": gb(rep_b, trn),
        "single_core": {"value": round(1.0 / per_byte_1 / 1e9, 6), "encode_gbps": gb(enc_b, te1),
                        "repair_gbps": gb(rep_b, tr1), "encode_s": round(te1, 3), "repair_s": round(tr1, 3)},
    }
    if il16 is not None:
        groups = (sample // 16) * 16
        rb16 = groups * k * elem
        res["repair_interleave16_gbps"] = gb(rb16, il16)
        res["value_interleave16"] = round(1.0 / (ten / enc_b + il16 / rb16) / 1e9, 6)
    return res


def rocprof_traffic(workload):
    """HBM bytes per dispatch of each kernel, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by what they read for a copy of known size)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None


def run_plain(name, args):
    """One step = encode of every message + repair of every corrupted codeword, with all
    inputs resident in HBM; the corrupted input is restored by a device copy before each
    step, outside the timed bracket."""
    from libffrs_amd import capi
    from libffrs_amd import workloads as W

    wl = dict(WORKLOADS[name])
    if args.codewords:
        wl["N"] = args.codewords
    N, n, ecc = wl["N"], wl["n"], wl["ecc"]
    k = n - ecc
    rs256 = wl["field"] == "rs256"
    elem = 1 if rs256 else 2
    t0 = time.perf_counter()
    if rs256:
        msgs, counts, s = W.rs256_plain(N, ecc, wl["max_err"], seed=wl["seed"], uncorrectable_frac=wl["unc"])
    else:
        msgs, counts, s = W.rsi16_plain(N, n, ecc, wl["max_err"], seed=wl["seed"])
    log(f"generated {N} messages in {time.perf_counter() - t0:.1f} s")

    ctx = capi.Context(0)
    dev_name = ctx.info()[0]
    if rs256:
        codec = capi.RS256Codec(ctx, ecc)
        encode = lambda dm, dp: codec.encode_device(k, dm, N, dp)  # noqa: E731
        repair = lambda dm, dp, ds: codec.repair_device(k, dm, dp, N, ds)  # noqa: E731
    else:
        codec = capi.RSi16Codec(ctx, n, ecc)
        encode = lambda dm, dp: codec.encode_device(dm, N, dp, True)  # noqa: E731
        repair = lambda dm, dp, ds: codec.repair_device(dm, dp, N, ds)  # noqa: E731
    mb, pb = N * k * elem, N * ecc * elem
    d_msg, d_par = ctx.alloc(mb), ctx.alloc(pb)
    d_cm, d_cp = ctx.alloc(mb), ctx.alloc(pb)     # pristine corrupted codewords
    d_wm, d_wp = ctx.alloc(mb), ctx.alloc(pb)     # working copy repaired in place
    d_st = ctx.alloc(N)
    d_msg.upload(msgs)
    encode(d_msg, d_par)
    par = np.empty((N, ecc), msgs.dtype)
    d_par.download(par)
    cws = np.concatenate([msgs, par], 1)
    t0 = time.perf_counter()
    (W.rs256_corrupt if rs256 else W.rsi16_corrupt)(cws, counts, s)
    log(f"corrupted codewords in {time.perf_counter() - t0:.1f} s")
    d_cm.upload(np.ascontiguousarray(cws[:, :k]))
    d_cp.upload(np.ascontiguousarray(cws[:, k:]))

    def step(instrumented):
        d_wm.copy_from(d_cm)
        d_wp.copy_from(d_cp)
        ctx.sync()
        t = time.perf_counter()
        encode(d_msg, d_par)
        ke = ctx.kernel_times() if instrumented else None
        repair(d_wm, d_wp, d_st)
        kr = ctx.kernel_times() if instrumented else None
        ctx.sync()
        return time.perf_counter() - t, ke, kr

    for _ in range(args.warmup):
        step(False)
    # timed steps: wall clock between two stream synchronisations
    walls = [step(False)[0] for _ in range(args.steps)]
    # the same steps again with a HIP event after every kernel, for the per-kernel split
    ctx.kernel_timing(True)
    kes, krs = [], []
    for _ in range(max(args.steps, 1)):
        _, ke, kr = step(True)
        kes.append(ke)
        krs.append(kr)
    ctx.kernel_timing(False)
    st = np.empty(N, np.uint8)
    d_st.download(st)
    ms_step = 1e3 * float(np.mean(walls))
    msg_bytes = N * k * elem
    value = msg_bytes / (ms_step * 1e-3) / 1e9
    if rs256:
        enc_names = ["rs256_encode_kernel"]
        # ecc 2, 4, 8, 16, 32 run BM inside the Chien kernel; other ecc <= 32 launch it on its own
        rep_names = (["rs256_synd_kernel", "rs256_chien16_kernel"] if ecc in (2, 4, 8, 16, 32)
                     else ["rs256_synd_kernel", "rs256_bm_kernel", "rs256_chien16_kernel"] if ecc <= 32
                     else ["rs256_repair_kernel"])
    else:
        # the int8 MFMA contraction where it applies (these shapes have no mix constant of 65536)
        enc_names = ["rsi16_encode_mfma_kernel" if ecc % 16 == 0 and ecc <= 256 and k <= 16384 else
                     "rsi16_encode_lane_kernel" if ecc <= 32 else
                     ("rsi16_encode_reg_kernel" if ecc <= 256 else "rsi16_encode_kernel")]
        # syndromes as an int8 MFMA contraction ahead of the register repair where its tables exist
        rep_names = ((["rsi16_synd_mfma_kernel"] if ecc in (16, 32, 64) and k <= 16384 else []) +
                     ["rsi16_repair_reg_kernel" if ecc <= 256 else "rsi16_repair_kernel"])
    ke_ms = np.mean(np.array(kes), axis=0)
    kr_ms = np.mean(np.array(krs), axis=0)
    if len(ke_ms) != len(enc_names) or len(kr_ms) != len(rep_names):
        raise RuntimeError(f"unexpected kernel count: encode {len(ke_ms)}, repair {len(kr_ms)}")
    te_ms, tr_ms = float(ke_ms.sum()), float(kr_ms.sum())
    # algorithmic bytes per launch: encode reads the message and writes the parity; repair
    # reads the codeword and writes one status byte (corrections are sparse and left out)
    enc_bytes = N * (k + ecc) * elem
    rep_bytes = N * n * elem + N
    traffic = rocprof_traffic(name)
    prefix = wl["field"]
    launches = {
        prefix + "_encode": {"kernels": enc_names, "ms": te_ms, "bytes": enc_bytes},
        prefix + "_repair": {"kernels": rep_names, "ms": tr_ms, "bytes": rep_bytes},
    }
    dom = max(launches, key=lambda x: launches[x]["ms"])
    d = launches[dom]
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9
    dom_traffic = None
    if traffic and all(x in traffic for x in d["kernels"]):
        dom_traffic = int(sum(traffic[x]["bytes"] for x in d["kernels"]))
    kernels = {}
    for names, times in ((enc_names, ke_ms), (rep_names, kr_ms)):
        for x, t in zip(names, times):
            kernels[x] = {"ms": round(float(t), 4)}
            if traffic and x in traffic:
                kernels[x]["traffic"] = int(traffic[x]["bytes"])
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if rs256 else "u32 (GF(65537) residues of u16 symbols)",
        "data": f"synthetic (SplitMix64 stream, seed {wl['seed']})",
        "config": {"workload": name, "codewords": N, "block_len": n, "ecc_len": ecc, "message_len": k,
                   "errors": wl["errors"], "device": dev_name},
        "roofline": {"bound": "hbm", "kernel": dom + " (" + " + ".join(d["kernels"]) + ")",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": dom_traffic},
        "launches": {x: {"ms": round(v["ms"], 4), "alg_bytes": v["bytes"],
                         "alg_gbps": round(v["bytes"] / v["ms"] / 1e6, 2),
                         "msg_gbps": round(msg_bytes / v["ms"] / 1e6, 2)} for x, v in launches.items()},
        "kernels": kernels,
        "status_counts": {int(v): int(c) for v, c in zip(*np.unique(st, return_counts=True))},
    }
    for b in (d_msg, d_par, d_cm, d_cp, d_wm, d_wp, d_st):
        b.free()
    codec.close()
    ctx.close()
    return result


def run_mixed(args):
    """Mixed-size batch: 200,000 RS256 codewords (ecc_len uniform in {2,4,8,16,32,64},
    message length uniform in [1, 255 - ecc_len], errors uniform in [0, ecc_len/2]) and
    50,000 RSi16 codewords (block_len in {64,256,1024}, ecc_len in {8,16,32}, errors in
    [0, ecc_len/2]). The fused path runs one persistent work-queue launch per field and
    stage, replayed from HIP graphs; the plain path runs the per-size entry points once per
    distinct (field, block_len, ecc_len), issued from the host and also replayed from one
    captured graph (its best case). Both paths are checked to give identical results."""
    from libffrs_amd import capi
    from libffrs_amd import workloads as W

    n_rs, n_ri = args.mixed_rs256, args.mixed_rsi16
    t0 = time.perf_counter()
    eccs, lens, counts, s = W.rs256_mixed(n_rs, seed=4)
    msg_all = s.bytes(int(lens.sum()))
    starts = np.concatenate([[0], np.cumsum(lens)])  # item starts in generation order
    # group-contiguous layout: groups of equal (ecc_len, message_len) stored back to back
    order = np.lexsort((lens, eccs))
    eccs, lens, counts = eccs[order], lens[order], counts[order]
    msgs = [msg_all[starts[o]:starts[o + 1]] for o in order]
    moff = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64)
    eoff = np.concatenate([[0], np.cumsum(eccs)[:-1]]).astype(np.uint64)
    hm = np.concatenate(msgs).astype(np.uint8)
    keys = np.stack([eccs, lens], 1)
    gstart = np.concatenate([[0], np.nonzero((keys[1:] != keys[:-1]).any(1))[0] + 1, [n_rs]])
    ns_, eccs_i, counts_i, s2 = W.rsi16_mixed(n_ri, seed=5)
    order_i = np.lexsort((eccs_i, ns_))
    ns_, eccs_i, counts_i = ns_[order_i], eccs_i[order_i], counts_i[order_i]
    ks_i = ns_ - eccs_i
    hm_i = s2.u16(int(ks_i.sum()))
    moff_i = np.concatenate([[0], np.cumsum(ks_i)[:-1]]).astype(np.uint64)
    eoff_i = np.concatenate([[0], np.cumsum(eccs_i)[:-1]]).astype(np.uint64)
    keys_i = np.stack([ns_, eccs_i], 1)
    gstart_i = np.concatenate([[0], np.nonzero((keys_i[1:] != keys_i[:-1]).any(1))[0] + 1, [n_ri]])
    log(f"mixed batch generated in {time.perf_counter() - t0:.1f} s: {len(gstart) - 1} RS256 and "
        f"{len(gstart_i) - 1} RSi16 size classes")

    ctx = capi.Context(0)
    plan = capi.RS256Plan(ctx, eccs, lens, moff, eoff, graph=not args.no_graph)
    plan_i = capi.RSi16Plan(ctx, ns_, eccs_i, moff_i, eoff_i, graph=not args.no_graph)
    tuned = {}
    if not args.no_tune:
        # per-class kernel shapes from the autotuner (its cache, tuning the classes it lacks)
        from libffrs_amd import autotune
        t1 = time.perf_counter()
        tuned = {"rs256": {str(e): r["mode"] for e, r in autotune.apply_plan(plan).items()},
                 "rsi16": {f"{n},{e}": [r["l_encode"], r["g_repair"]] for (n, e), r in autotune.apply_plan(plan_i).items()}}
        log(f"plan shapes from the autotuner in {time.perf_counter() - t1:.1f} s: {tuned}")
    d_m, d_p = ctx.alloc(hm.nbytes), ctx.alloc(int(eccs.sum()))
    d_mi, d_pi = ctx.alloc(hm_i.nbytes), ctx.alloc(int(eccs_i.sum()) * 2)
    d_m.upload(hm)
    d_mi.upload(hm_i)
    plan.encode(d_m, d_p)
    plan_i.encode(d_mi, d_pi)
    hp = np.empty(int(eccs.sum()), np.uint8)
    hp_i = np.empty(int(eccs_i.sum()), np.uint16)
    d_p.download(hp)
    d_pi.download(hp_i)
    # corrupt group by group (rows of one group share their length)
    cm, cp = hm.copy(), hp.copy()
    for a, b in zip(gstart[:-1], gstart[1:]):
        k, e = int(lens[a]), int(eccs[a])
        m0, p0 = int(moff[a]), int(eoff[a])
        cw = np.concatenate([cm[m0:m0 + (b - a) * k].reshape(-1, k), cp[p0:p0 + (b - a) * e].reshape(-1, e)], 1)
        W.rs256_corrupt(cw, counts[a:b], s)
        cm[m0:m0 + (b - a) * k] = cw[:, :k].ravel()
        cp[p0:p0 + (b - a) * e] = cw[:, k:].ravel()
    cm_i, cp_i = hm_i.copy(), hp_i.copy()
    for a, b in zip(gstart_i[:-1], gstart_i[1:]):
        k, e = int(ks_i[a]), int(eccs_i[a])
        m0, p0 = int(moff_i[a]), int(eoff_i[a])
        cw = np.concatenate([cm_i[m0:m0 + (b - a) * k].reshape(-1, k),
                             cp_i[p0:p0 + (b - a) * e].reshape(-1, e)], 1)
        W.rsi16_corrupt(cw, counts_i[a:b], s2)
        cm_i[m0:m0 + (b - a) * k] = cw[:, :k].ravel()
        cp_i[p0:p0 + (b - a) * e] = cw[:, k:].ravel()
    d_cm, d_cp, d_cmi, d_cpi = ctx.alloc(cm.nbytes), ctx.alloc(cp.nbytes), ctx.alloc(cm_i.nbytes), ctx.alloc(cp_i.nbytes)
    d_cm.upload(cm)
    d_cp.upload(cp)
    d_cmi.upload(cm_i)
    d_cpi.upload(cp_i)
    d_wm, d_wp, d_wmi, d_wpi = ctx.alloc(cm.nbytes), ctx.alloc(cp.nbytes), ctx.alloc(cm_i.nbytes), ctx.alloc(cp_i.nbytes)
    d_st, d_sti = ctx.alloc(n_rs), ctx.alloc(n_ri)

    codecs = {int(e): capi.RS256Codec(ctx, int(e)) for e in np.unique(eccs)}
    codecs_i = {(int(n), int(e)): capi.RSi16Codec(ctx, int(n), int(e)) for n, e in np.unique(keys_i, axis=0)}

    def restore():
        d_wm.copy_from(d_cm)
        d_wp.copy_from(d_cp)
        d_wmi.copy_from(d_cmi)
        d_wpi.copy_from(d_cpi)
        ctx.sync()

    def fused_encode():
        plan.encode(d_m, d_p)
        plan_i.encode(d_mi, d_pi)

    def fused_repair():
        plan.repair(d_wm, d_wp, d_st)
        plan_i.repair(d_wmi, d_wpi, d_sti)

    # the same two plans with the RSi16 one on a second context (stream), so that the two
    # persistent launches run side by side; timed on the host around both streams
    ctx2 = capi.Context(0)
    plan_i2 = capi.RSi16Plan(ctx2, ns_, eccs_i, moff_i, eoff_i, graph=not args.no_graph)

    def fused2(repair):
        if repair:
            plan.repair(d_wm, d_wp, d_st)
            plan_i2.repair(d_wmi, d_wpi, d_sti)
        else:
            plan.encode(d_m, d_p)
            plan_i2.encode(d_mi, d_pi)
        ctx.sync()
        ctx2.sync()

    def plain_encode():
        for a, b in zip(gstart[:-1], gstart[1:]):
            k, e = int(lens[a]), int(eccs[a])
            codecs[e].encode_device(k, d_m.view(int(moff[a]), (b - a) * k), int(b - a),
                                    d_p.view(int(eoff[a]), (b - a) * e))
        for a, b in zip(gstart_i[:-1], gstart_i[1:]):
            k, e = int(ks_i[a]), int(eccs_i[a])
            codecs_i[(int(ns_[a]), e)].encode_device(d_mi.view(2 * int(moff_i[a]), 2 * (b - a) * k), int(b - a),
                                                     d_pi.view(2 * int(eoff_i[a]), 2 * (b - a) * e))

    def plain_repair():
        for a, b in zip(gstart[:-1], gstart[1:]):
            k, e = int(lens[a]), int(eccs[a])
            codecs[e].repair_device(k, d_wm.view(int(moff[a]), (b - a) * k), d_wp.view(int(eoff[a]), (b - a) * e),
                                    int(b - a), d_st.view(int(a), int(b - a)))
        for a, b in zip(gstart_i[:-1], gstart_i[1:]):
            k, e = int(ks_i[a]), int(eccs_i[a])
            codecs_i[(int(ns_[a]), e)].repair_device(d_wmi.view(2 * int(moff_i[a]), 2 * (b - a) * k),
                                                     d_wpi.view(2 * int(eoff_i[a]), 2 * (b - a) * e), int(b - a),
                                                     d_sti.view(int(a), int(b - a)))

    def results():
        out = [np.empty_like(cm), np.empty_like(cp), np.empty(n_rs, np.uint8), np.empty_like(cm_i),
               np.empty_like(cp_i), np.empty(n_ri, np.uint8)]
        for b, o in zip((d_wm, d_wp, d_st, d_wmi, d_wpi, d_sti), out):
            b.download(o)
        return out

    def timed(fn, reps, before=None):
        ms = []
        for _ in range(reps):
            if before:
                before()
            ctx.sync()
            t = time.perf_counter()
            ctx.timer_begin()
            fn()
            gpu = ctx.timer_end()
            ms.append((1e3 * (time.perf_counter() - t), gpu))
        return float(np.median([m[0] for m in ms])), float(np.median([m[1] for m in ms]))

    # correctness: fused and plain give the same parity, bytes and statuses
    fused_encode()
    ctx.sync()
    restore()
    fused_repair()
    ctx.sync()
    r_fused = results()
    par_fused = np.empty_like(hp)
    d_p.download(par_fused)
    plain_encode()
    restore()
    plain_repair()
    ctx.sync()
    r_plain = results()
    par_plain = np.empty_like(hp)
    d_p.download(par_plain)
    same = bool((par_fused == par_plain).all() and all((x == y).all() for x, y in zip(r_fused, r_plain)))
    if not same:
        raise RuntimeError("fused and plain mixed-batch results differ")

    reps = max(args.steps, 3)
    for _ in range(args.warmup):
        fused_encode()
        restore()
        fused_repair()
    fe = timed(fused_encode, reps)
    fr = timed(fused_repair, reps, restore)
    fused2(False)
    restore()
    fused2(True)
    r_two = results()
    if not all((x == y).all() for x, y in zip(r_two, r_fused)):
        raise RuntimeError("two-stream fused results differ")
    fe2 = timed(lambda: fused2(False), reps)
    fr2 = timed(lambda: fused2(True), reps, restore)
    pe = timed(plain_encode, reps)
    pr = timed(plain_repair, reps, restore)
    pge = pgr = (float("nan"), float("nan"))
    if not args.no_graph:
        # the plain launches replayed from one captured graph: no host issue cost
        ctx.capture_begin()
        plain_encode()
        g_pe = ctx.capture_end()
        ctx.capture_begin()
        plain_repair()
        g_pr = ctx.capture_end()
        pge = timed(g_pe.launch, reps)
        pgr = timed(g_pr.launch, reps, restore)
        g_pe.close()
        g_pr.close()

    msg_bytes = hm.nbytes + hm_i.nbytes
    st_f, sti_f = r_fused[2], r_fused[5]

    def gbps(ms):
        return round(msg_bytes / (ms * 1e-3) / 1e9, 3)

    res = {
        "workload": "mixed (200,000 RS256 + 50,000 RSi16 codewords)",
        "codewords": {"rs256": n_rs, "rsi16": n_ri},
        "size_classes": {"rs256": len(gstart) - 1, "rsi16": len(gstart_i) - 1},
        "message_bytes": msg_bytes,
        "fused": {"encode_ms": round(fe[1], 4), "repair_ms": round(fr[1], 4),
                  "encode_gbps": gbps(fe[1]), "repair_gbps": gbps(fr[1])},
        "fused_two_streams": {"encode_ms": round(fe2[0], 4), "repair_ms": round(fr2[0], 4),
                              "encode_gbps": gbps(fe2[0]), "repair_gbps": gbps(fr2[0]),
                              "timing": "host wall clock around both streams"},
        "plain_host_issued": {"encode_ms": round(pe[0], 4), "repair_ms": round(pr[0], 4),
                              "encode_gpu_ms": round(pe[1], 4), "repair_gpu_ms": round(pr[1], 4)},
        "plain_graph": ({"encode_ms": round(pge[1], 4), "repair_ms": round(pgr[1], 4)}
                        if not args.no_graph else None),
        "repair_speedup_vs_plain": round(pr[0] / fr[1], 2),
        "repair_speedup_two_streams_vs_plain": round(pr[0] / fr2[0], 2),
        "repair_speedup_vs_plain_graph": round(pgr[1] / fr[1], 2) if not args.no_graph else None,
        "encode_speedup_vs_plain_graph": round(pge[1] / fe[1], 2) if not args.no_graph else None,
        "graphs": not args.no_graph,
        "tuned_shapes": tuned or None,
        "identical_results": same,
        "status_counts": {"rs256": {int(v): int(c) for v, c in zip(*np.unique(st_f, return_counts=True))},
                          "rsi16": {int(v): int(c) for v, c in zip(*np.unique(sti_f, return_counts=True))}},
    }
    for c in list(codecs.values()) + list(codecs_i.values()):
        c.close()
    plan.close()
    plan_i.close()
    plan_i2.close()
    for b in (d_m, d_p, d_mi, d_pi, d_cm, d_cp, d_cmi, d_cpi, d_wm, d_wp, d_wmi, d_wpi, d_st, d_sti):
        b.free()
    ctx2.close()
    ctx.close()
    return res


def run_serving(args):
    """Serving loop: Poisson arrivals of par-style requests, 70% encode and 30% verify+repair,
    files log-uniform in 4 KiB-64 MiB split into RSi16(1024, 128) blocks (ecc 128: 12.5% parity,
    the legal neighbour of 10%), repair requests corrupted at 0-5% of their symbols (half single
    errors, half 8-64 symbol bursts). Requests are batched continuously by
    libffrs_amd.serving.BlockServer; latency runs from submission to the moment the last
    block's results are back in host memory. The same request set is replayed at
    args.serve_rate and at the rates of args.serve_sweep (offered load rising until the server
    saturates); the main fields are those of args.serve_rate."""
    from libffrs_amd import workloads as W
    from libffrs_amd.serving import BlockServer

    n_req = args.serve_requests
    t_arr, kinds, sizes, s = W.serving_requests(n_req, 1.0, seed=6)  # arrival times at 1 request/s
    t0 = time.perf_counter()
    datas = [s.bytes(int(sz)) for sz in sizes]
    srv = BlockServer(block_len=1024, ecc_len=128, slots=args.serve_slots, slot_blocks=args.serve_slot_blocks,
                      piece_blocks=2048)
    # repair requests: encode their files first (not timed), then corrupt codewords
    payload = []
    originals = {}
    for i, (kind, d) in enumerate(zip(kinds, datas)):
        if kind == "encode":
            payload.append((kind, d, None))
            continue
        r = srv.submit("encode", d)
        r.done.wait()
        msgs, par = r.msgs.copy(), r.parity.copy()
        originals[i] = r.msgs.copy()
        W.serving_corrupt(msgs, par, s)
        payload.append((kind, msgs.reshape(-1).view(np.uint8), par))
    log(f"serving: {n_req} requests prepared in {time.perf_counter() - t0:.1f} s ({sizes.sum() / 1e9:.2f} GB)")

    def run_at(rate):
        srv.batches.clear()
        reqs = []
        start = time.perf_counter()
        for i, (kind, d, par) in enumerate(payload):
            delay = start + t_arr[i] / rate - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            reqs.append(srv.submit(kind, d, par))
        for r in reqs:
            r.done.wait()
        end = max(r.t_done for r in reqs)
        lat = np.array([r.latency for r in reqs]) * 1e3
        total = sum(r.nbytes for r in reqs)
        restored = sum(int((reqs[i].msgs == originals[i]).all(axis=1).sum()) for i in originals)
        rep_blocks = sum(reqs[i].blocks for i in originals)
        ok_blocks = sum(int(np.isin(reqs[i].status, (0, 2)).sum()) for i in originals)
        return {
            "rate_per_s": rate,
            "offered_gbps": round(float(sizes.mean()) * rate / 1e9, 3),
            "bytes": int(total),
            "seconds": round(end - start, 3),
            "gbps": round(total / (end - start) / 1e9, 3),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3),
                           "max": round(float(lat.max()), 3), "mean": round(float(lat.mean()), 3)},
            "batches": len(srv.batches),
            "batches_with_several_requests": sum(1 for b in srv.batches if b[1] > 1),
            "repair_blocks": rep_blocks,
            "repair_blocks_ok_status": ok_blocks,
            "repair_blocks_restored": restored,
        }

    res = {"workload": "serving (Poisson arrivals, 70% encode / 30% verify+repair, RSi16(1024, 128) blocks)",
           "requests": n_req, "slots": args.serve_slots, "hbm_free_bytes": srv.hbm_free,
           "hbm_budget_bytes": srv.hbm_budget}
    res.update(run_at(args.serve_rate))
    sweep = [float(x) for x in args.serve_sweep.split(",") if x] if args.serve_sweep else []
    if sweep:
        res["sweep"] = []
        for rate in sweep:
            r = run_at(rate)
            res["sweep"].append({k_: r[k_] for k_ in ("rate_per_s", "offered_gbps", "gbps", "latency_ms", "batches",
                                                      "batches_with_several_requests")})
            log(f"serving at {rate}/s: offered {r['offered_gbps']} GB/s, achieved {r['gbps']} GB/s, "
                f"p99 {r['latency_ms']['p99']} ms")
    srv.close()
    return res


def run_drop_in(args):
    """Per-call latency of the drop-in API on the reference's own calling pattern, one codeword
    per call (host buffers in, results back in host memory): RS256(ecc_len=32).encode / .repair
    of 223-byte messages and RSi16(256, 32).encode / .repair, next to the reference's CPU build
    doing the same calls. Microseconds per call, median and p99 over `calls` calls."""
    import libffrs_amd as L
    from libffrs_amd import workloads as W
    ref = load_reference_cpu()
    calls = args.drop_in_calls

    def time_calls(fn, items):
        for it in items[:10]:
            fn(*it)
        ts = []
        for it in items:
            t = time.perf_counter()
            fn(*it)
            ts.append(time.perf_counter() - t)
        ts = np.array(ts) * 1e6
        return {"median_us": round(float(np.median(ts)), 2), "p99_us": round(float(np.percentile(ts, 99)), 2)}

    out = {"calls": calls}
    msgs, counts, s = W.rs256_plain(calls, 32, 16, seed=21)
    gpu_rs = L.RS256(ecc_len=32)
    par = np.frombuffer(bytes(gpu_rs.encode(msgs.tobytes())), np.uint8).reshape(calls, 32)
    cws = np.concatenate([msgs, par], 1)
    W.rs256_corrupt(cws, counts, s)

    def rs_items():
        return [(bytearray(r[:223].tobytes()), bytearray(r[223:].tobytes())) for r in cws]

    enc_items = [(m.tobytes(),) for m in msgs]
    out["rs256_ecc32"] = {"gpu_encode": time_calls(gpu_rs.encode, enc_items),
                          "gpu_repair": time_calls(gpu_rs.repair, rs_items())}
    if ref is not None:
        cpu_rs = ref.RS256(ecc_len=32)
        out["rs256_ecc32"]["reference_cpu_encode"] = time_calls(cpu_rs.encode, enc_items)
        out["rs256_ecc32"]["reference_cpu_repair"] = time_calls(cpu_rs.repair, rs_items())
    msgs, counts, s = W.rsi16_plain(calls, 256, 32, 16, seed=22)
    gpu_ri = L.RSi16(256, 32)
    par = np.frombuffer(bytes(gpu_ri.encode(msgs.tobytes())), np.uint16).reshape(calls, 32)
    cws = np.concatenate([msgs, par], 1)
    W.rsi16_corrupt(cws, counts, s)

    def ri_items():
        return [(bytearray(r[:224].tobytes()), bytearray(r[224:].tobytes())) for r in cws]

    enc_items = [(m.tobytes(),) for m in msgs]
    out["rsi16_256_32"] = {"gpu_encode": time_calls(gpu_ri.encode, enc_items),
                           "gpu_repair": time_calls(gpu_ri.repair, ri_items())}
    if ref is not None:
        cpu_ri = ref.RSi16(256, 32)
        out["rsi16_256_32"]["reference_cpu_encode"] = time_calls(cpu_ri.encode, enc_items)
        out["rsi16_256_32"]["reference_cpu_repair"] = time_calls(cpu_ri.repair, ri_items())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="rs256_plain_scale", choices=sorted(WORKLOADS))
    ap.add_argument("--codewords", type=int, default=0, help="override the workload's codeword count")
    ap.add_argument("--cpu-repair-sample", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed", action="store_true", help="run only the mixed-size batch and print its result")
    ap.add_argument("--no-mixed", action="store_true", help="leave the mixed-size batch out of the bench line")
    ap.add_argument("--no-graph", action="store_true", help="mixed batch without HIP graphs (for tracing tools)")
    ap.add_argument("--no-tune", action="store_true", help="mixed batch with the plans' default shapes")
    ap.add_argument("--serving", action="store_true", help="run only the serving loop and print its result")
    ap.add_argument("--no-serving", action="store_true", help="leave the serving loop out of the bench line")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=800.0)
    ap.add_argument("--serve-sweep", default="3200,12800,51200",
                    help="further request rates (per second) the serving request set is replayed at")
    ap.add_argument("--drop-in-calls", type=int, default=300)
    ap.add_argument("--no-drop-in", action="store_true", help="leave the per-call drop-in latencies out")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-slot-blocks", type=int, default=16384)
    ap.add_argument("--mixed-rs256", type=int, default=200000)
    ap.add_argument("--mixed-rsi16", type=int, default=50000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported by this build "
              "(codewords are independent; a multi-GPU run would shard them with no collectives)",
              file=sys.stderr)
        sys.exit(2)
    ensure_built()
    if args.mixed:
        print(json.dumps(run_mixed(args)), flush=True)
        return
    if args.serving:
        print(json.dumps(run_serving(args)), flush=True)
        return
    cpu = None
    if not args.no_cpu_baseline:
        # first, while no GPU state exists (the baseline forks worker processes)
        wl = dict(WORKLOADS[args.workload])
        from libffrs_amd import workloads as W
        N = args.codewords or wl["N"]
        log("timing the CPU baseline ...")
        if wl["field"] == "rs256":
            msgs_cpu = W.rs256_plain(N, wl["ecc"], wl["max_err"], seed=wl["seed"], uncorrectable_frac=wl["unc"])[0]
        else:
            msgs_cpu = W.rsi16_plain(N, wl["n"], wl["ecc"], wl["max_err"], seed=wl["seed"])[0]
        cpu = cpu_baseline(wl, msgs_cpu, min(N, args.cpu_repair_sample or wl["cpu_repair"]))
        del msgs_cpu
    res = run_plain(args.workload, args)
    if cpu is not None:
        res["cpu_baseline"] = cpu
    if not args.no_mixed:
        log("running the mixed-size batch ...")
        res["mixed"] = run_mixed(args)
    if not args.no_serving:
        log("running the serving loop ...")
        res["serving"] = run_serving(args)
    if not args.no_drop_in:
        log("timing drop-in calls ...")
        res["drop_in"] = run_drop_in(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
