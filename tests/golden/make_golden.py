"""Write the golden fixtures of tests/golden/*.npz from the reference's own code.

Runs only where the reference source is present: it imports the reference's
pybind11 extension as compiled from its sources by `make -C oracle ref` into
oracle/_ref/ (recipe: oracle/Makefile). The fixtures hold input digests and
per-codeword outputs (parity CRCs, repair status, repaired-codeword CRCs); the
small edge sets hold their inputs and outputs explicitly.

    python tests/golden/make_golden.py
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import cases as C  # noqa: E402
from libffrs_amd import workloads as W  # noqa: E402


def load_reference():
    import sysconfig
    path = os.path.join(ROOT, "oracle", "_ref", "libffrs" + sysconfig.get_config_var("EXT_SUFFIX"))
    if not os.path.exists(path):
        raise SystemExit("build the reference first: make -C oracle ref")
    spec = importlib.util.spec_from_file_location("libffrs", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def ref_repair_rs256(ref, cws, k):
    status = np.zeros(len(cws), np.uint8)
    out = cws.copy()
    for i in range(len(cws)):
        m = bytearray(cws[i, :k].tobytes())
        e = bytearray(cws[i, k:].tobytes())
        status[i] = 1 if ref.repair(m, e) else 0
        out[i, :k] = np.frombuffer(m, np.uint8)
        out[i, k:] = np.frombuffer(e, np.uint8)
    return status, out


def ref_repair_rsi16(ref, cws, k, positions=None):
    status = np.zeros(len(cws), np.uint8)
    out = cws.copy()
    for i in range(len(cws)):
        m = bytearray(cws[i, :k].tobytes())
        e = bytearray(cws[i, k:].tobytes())
        try:
            if positions is None:
                status[i] = int(ref.repair(m, e))
            else:
                status[i] = int(ref.repair(m, e, [int(p) for p in positions[i] if p >= 0]))
        except RuntimeError:
            status[i] = 255
            continue
        out[i, :k] = np.frombuffer(m, np.uint16)
        out[i, k:] = np.frombuffer(e, np.uint16)
    return status, out


def main():
    L = load_reference()
    out = {}

    # ---------------------------------------------------------------- RS256 plain
    for name in C.RS256_PLAIN_CASES:
        ecc, msgs, counts, s = C.rs256_plain_case(name)
        k = 255 - ecc
        ref = L.RS256(ecc_len=ecc)
        par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint8).reshape(-1, ecc)
        cws = np.concatenate([msgs, par], 1)
        W.rs256_corrupt(cws, counts, s)
        st, rep = ref_repair_rs256(ref, cws, k)
        out[f"rs256_{name}"] = dict(in_sha=C.sha(msgs, counts), corrupt_sha=C.sha(cws),
                                   parity_crc=C.row_crc(par), status=st, repaired_crc=C.row_crc(rep))
        print(name, "status true:", int(st.sum()), "/", len(st))

    # ---------------------------------------------------------------- RS256 shapes
    for (bl, ecc) in C.RS256_SHAPES:
        k = bl - ecc
        buf, rep_msgs, counts, s = C.rs256_shape_case(bl, ecc)
        ref = L.RS256(block_len=bl, ecc_len=ecc)
        par_all = np.frombuffer(ref.encode(buf.tobytes()), np.uint8)
        par = np.frombuffer(ref.encode(rep_msgs.tobytes()), np.uint8).reshape(-1, ecc)
        cws = np.concatenate([rep_msgs, par], 1)
        W.rs256_corrupt(cws, counts, s)
        st, rep = ref_repair_rs256(ref, cws, k)
        synd = np.frombuffer(ref._synds(cws[0].tobytes()), np.uint8)
        out[f"rs256_shape_{bl}_{ecc}"] = dict(
            in_sha=C.sha(buf, rep_msgs, counts), parity_all=par_all, corrupt_sha=C.sha(cws), status=st,
            repaired_crc=C.row_crc(rep), synds0=synd, generator=np.frombuffer(ref.generator, np.uint8))

    # ---------------------------------------------------------------- RS256 mixed
    eccs, lens, counts, msgs, s = C.rs256_mixed_case()
    refs = {}
    par_crc = np.zeros(len(eccs), np.uint32)
    st = np.zeros(len(eccs), np.uint8)
    rep_crc = np.zeros(len(eccs), np.uint32)
    corrupt = []
    for i, (e, l) in enumerate(zip(eccs, lens)):
        e, l = int(e), int(l)
        ref = refs.setdefault((l, e), L.RS256(block_len=l + e, ecc_len=e))
        par = np.frombuffer(ref.encode(msgs[i].tobytes()), np.uint8)
        par_crc[i] = C.row_crc(par[None])[0]
        cw = np.concatenate([msgs[i], par])[None].copy()
        W.rs256_corrupt(cw, counts[i:i + 1], s)
        corrupt.append(cw[0].copy())
        s1, r1 = ref_repair_rs256(ref, cw, l)
        st[i] = s1[0]
        rep_crc[i] = C.row_crc(r1)[0]
    out["rs256_mixed"] = dict(in_sha=C.sha(eccs, lens, counts, *msgs), corrupt_sha=C.sha(*corrupt),
                              parity_crc=par_crc, status=st, repaired_crc=rep_crc)

    # ---------------------------------------------------------------- RSi16 plain
    for name in C.RSI16_PLAIN_CASES:
        (n, ecc), msgs, counts, s = C.rsi16_plain_case(name)
        k = n - ecc
        ref = L.RSi16(n, ecc)
        par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
        refs_ = L.RSi16(n, ecc, simd_x4=False, simd_x8=False, simd_x16=False)
        par_s = np.frombuffer(refs_.encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
        cws = np.concatenate([msgs, par], 1)
        W.rsi16_corrupt(cws, counts, s)
        st, rep = ref_repair_rsi16(ref, cws, k)
        d = dict(in_sha=C.sha(msgs, counts), corrupt_sha=C.sha(cws), parity_crc=C.row_crc(par),
                 parity_scalar_crc=C.row_crc(par_s), status=st, repaired_crc=C.row_crc(rep),
                 root=np.array([ref.root]), ntt_len=np.array([ref.ntt_len]))
        # syndromes and Sugiyama outputs of the first 128 codewords (debug surface)
        syn = np.zeros((128, ecc), np.uint32)
        sug = np.zeros(128, np.uint32)
        for i in range(min(128, len(cws))):
            syn[i] = np.array(ref._synd(cws[i, :k].tobytes(), cws[i, k:].tobytes()), np.uint32)
            if syn[i].any():
                loc, ev = ref._sugiyama([int(v) for v in syn[i]])
                sug[i] = C.row_crc(np.array(list(loc) + list(ev), np.uint32)[None])[0]
        d["synds_head"] = syn
        d["sugiyama_crc_head"] = sug
        out[f"rsi16_{name}"] = d
        print(name, "status counts:", np.bincount(st, minlength=256)[[0, 2, 4, 255]])

    # ---------------------------------------------------------------- RSi16 mixed
    ns, eccs, counts, msgs, s = C.rsi16_mixed_case()
    par_crc = np.zeros(len(ns), np.uint32)
    st = np.zeros(len(ns), np.uint8)
    rep_crc = np.zeros(len(ns), np.uint32)
    corrupt = []
    refs = {}
    for i, (n, e) in enumerate(zip(ns, eccs)):
        n, e = int(n), int(e)
        ref = refs.setdefault((n, e), L.RSi16(n, e))
        par = np.frombuffer(ref.encode(msgs[i].tobytes()), np.uint16)
        par_crc[i] = C.row_crc(par[None])[0]
        cw = np.concatenate([msgs[i], par])[None].copy()
        W.rsi16_corrupt(cw, counts[i:i + 1], s)
        corrupt.append(cw[0].copy())
        s1, r1 = ref_repair_rsi16(ref, cw, n - e)
        st[i] = s1[0]
        rep_crc[i] = C.row_crc(r1)[0]
    out["rsi16_mixed"] = dict(in_sha=C.sha(ns, eccs, counts, *msgs), corrupt_sha=C.sha(*corrupt),
                              parity_crc=par_crc, status=st, repaired_crc=rep_crc)

    # ---------------------------------------------------------------- RSi16 edges
    n, ecc, k = 256, 32, 224
    ref = L.RSi16(n, ecc)
    # (a) messages whose true parity contains 65536 (stored as 0 by the API)
    s = W.Stream(13)
    trunc_msgs = []
    for _ in range(400):
        m = s.u16(4096 * k).reshape(4096, k)
        par = np.frombuffer(ref.encode(m.tobytes()), np.uint16).reshape(-1, ecc)
        for i in range(len(m)):
            cw_m = bytearray(m[i].tobytes())
            cw_e = bytearray(par[i].tobytes())
            if int(ref.repair(cw_m, cw_e)) != 0:   # a clean codeword that is not error-free
                trunc_msgs.append(m[i].copy())
        if len(trunc_msgs) >= 48:
            break
    trunc_msgs = np.array(trunc_msgs[:48], np.uint16)
    par = np.frombuffer(ref.encode(trunc_msgs.tobytes()), np.uint16).reshape(-1, ecc)
    cws = np.concatenate([trunc_msgs, par], 1)
    st, rep = ref_repair_rsi16(ref, cws, k)
    out["rsi16_edge_trunc"] = dict(msgs=trunc_msgs, parity=par, status=st, repaired_crc=C.row_crc(rep))
    print("truncation codewords:", len(trunc_msgs), "statuses", np.bincount(st))

    # (b) error patterns with last syndrome zero
    roots = []
    w = ref.root
    for i in range(ref.ntt_len):
        roots.append(pow(w, i, 65537))
    pats = C.rsi16_last_synd_zero(roots, n, ecc, 96)
    s = W.Stream(14)
    m = s.u16(len(pats) * k).reshape(len(pats), k)
    par = np.frombuffer(ref.encode(m.tobytes()), np.uint16).reshape(-1, ecc)
    cws = np.concatenate([m, par], 1).astype(np.int64)
    keep = []
    for i, (pos, val) in enumerate(pats):
        nv = (cws[i, pos] + val) % 65537
        if (nv == 65536).any():
            continue
        cws[i, pos] = nv
        keep.append(i)
    cws = cws[keep].astype(np.uint16)
    st, rep = ref_repair_rsi16(ref, cws, k)
    out["rsi16_edge_lastsynd"] = dict(cws=cws, status=st, repaired_crc=C.row_crc(rep))
    print("last-synd-zero codewords:", len(cws), "statuses", np.bincount(st))

    # (c) constant messages
    const = np.array([np.zeros(k), np.full(k, 65535), np.full(k, 1), np.arange(k) * 293 % 65536], np.uint16)
    par = np.frombuffer(ref.encode(const.tobytes()), np.uint16).reshape(-1, ecc)
    par_s = np.frombuffer(L.RSi16(n, ecc, simd_x4=False, simd_x8=False, simd_x16=False)
                          .encode(const.tobytes()), np.uint16).reshape(-1, ecc)
    out["rsi16_edge_const"] = dict(msgs=const, parity=par, parity_scalar=par_s)

    # (d) known-position repair: the corrupted positions plus up to 4 clean ones
    msgs, counts, s = C.rsi16_known_case()
    par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
    cws = np.concatenate([msgs, par], 1)
    pos = W.rsi16_corrupt(cws, counts, s)
    allpos = C.rsi16_known_positions(pos, s)
    st, rep = ref_repair_rsi16(ref, cws, k, positions=allpos)
    out["rsi16_known"] = dict(in_sha=C.sha(msgs, counts), corrupt_sha=C.sha(cws), positions=allpos, status=st,
                              repaired_crc=C.row_crc(rep))
    print("known-position statuses", np.bincount(st, minlength=5))

    out.update(radius_fixture(L))
    out.update(interleave_fixture(L))
    out.update(interleave_repair_fixture(L))
    out.update(reftest_fixture(L))
    out.update(divergence_fixture(L))
    out.update(ilsimd_fixture(L))
    out.update(ntt_fixture(L))
    out.update(circ16_fixture(L))
    out.update(circ16_planted_fixture(L))

    for key, d in out.items():
        np.savez_compressed(os.path.join(HERE, key + ".npz"), **d)
    total = sum(os.path.getsize(os.path.join(HERE, key + ".npz")) for key in out)
    print("wrote", len(out), "fixtures,", total, "bytes")


def radius_fixture(L):
    """(e) codewords with at most ecc/2 errors and exact parity on which the reference's
    repair does not restore the codeword, plus as many ordinary ones; the oracle only
    proposes candidates, the reference decides every expected output."""
    from oracle import oracle as O
    n, ecc, k = 256, 32, 224
    ref = L.RSi16(n, ecc)
    msgs, counts, s = C.rsi16_radius_case()
    par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
    orc = O.RSi16(n, ecc)
    cws = np.concatenate([msgs, par], 1)
    W.rsi16_corrupt(cws, counts, s)
    cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
    ost = orc.repair_batch(cm.copy(), cp.copy())
    cand = [i for i in np.nonzero(~np.isin(ost, (0, 2)))[0] if not orc.synds(msgs[i], par[i]).any()]
    rows = np.array(sorted(cand + list(range(0, 16 * 97, 97))), np.int64)
    st, rep = ref_repair_rsi16(ref, cws[rows], k)
    print("radius rows:", len(cand), "reference statuses", np.bincount(st, minlength=5))
    return {"rsi16_edge_radius": dict(cws=cws[rows], counts=counts[rows], status=st, repaired=rep)}


def radius1024_fixture(L):
    """(e) RSi16(1024, 128) blocks with at most 60 additive errors (ecc/2 = 64, so still within
    the radius where a parity symbol of 65536 was stored as 0) that the reference's repair
    leaves unrepaired, or reports repaired with other contents than were encoded, found by the
    reference itself over 160,000 blocks (all of the latter, 24 of the former), plus 8 ordinary ones; the reference decides every
    expected output."""
    n, ecc, k = 1024, 128, 896
    ref = L.RSi16(n, ecc)
    msgs, counts, s = W.rsi16_plain(160000, n, ecc, 60, seed=1024)
    par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
    cws = np.concatenate([msgs, par], 1)
    W.rsi16_corrupt(cws, counts, s)
    st, rep = ref_repair_rsi16(ref, cws, k)
    fail = np.nonzero(~np.isin(st, (0, 2)))[0]
    wrong = np.nonzero(np.isin(st, (0, 2)) & (rep != np.concatenate([msgs, par], 1)).any(axis=1))[0]
    rows = np.array(sorted(set(fail[:24].tolist()) | set(wrong.tolist()) | set(range(0, 8 * 9973, 9973))), np.int64)
    print("radius 1024 rows:", len(fail), "unrepaired,", len(wrong), "repaired to other contents, of", len(st),
          "; reference statuses", np.bincount(st[rows], minlength=5))
    return {"rsi16_edge_radius_1024": dict(cws=cws[rows], counts=counts[rows], status=st[rows],
                                           repaired_crc=C.row_crc(rep[rows]),
                                           encoded_crc=C.row_crc(np.concatenate([msgs, par], 1)[rows]))}


INTERLEAVE_CASES = ((64, 8, 3), (256, 32, 16), (256, 32, 5), (1024, 128, 2), (64, 16, 33))


def interleave_fixture(L):
    """(f) interleaved encode, RSi16(n, ecc, interleave): one interleaved block per call (the
    reference leaves the later blocks of a multi-block call uninitialised), with the
    host's SIMD words and with the scalar words."""
    out = {}
    s = W.Stream(17)
    for n, ecc, il in INTERLEAVE_CASES:
        k = n - ecc
        msgs = s.u16(6 * k * il).reshape(6, k * il)
        vec = L.RSi16(n, ecc, il)
        sca = L.RSi16(n, ecc, il, simd_x4=False, simd_x8=False, simd_x16=False)
        pv = np.stack([np.frombuffer(vec.encode(m.tobytes()), np.uint16) for m in msgs])
        ps = np.stack([np.frombuffer(sca.encode(m.tobytes()), np.uint16) for m in msgs])
        out[f"rsi16_interleave_{n}_{ecc}_{il}"] = dict(msgs=msgs, parity=pv, parity_scalar=ps)
    return out


IL_REPAIR_CASES = ((64, 8, 3), (256, 32, 5), (64, 16, 33), (1024, 128, 2))


def interleave_repair_fixture(L):
    """(f) interleaved repair with the scalar words (simd_x4/x8/x16 False), where the reference
    repairs the columns one after another (pyrsi16.hpp:391-453 with SIMD_W = 1): per trial the
    corrupted interleaved message and ecc, the error positions handed over (-1 padded; none
    in the first half of the trials), the returned status (255: the call raises) and both
    buffers after the call. Columns carry 0 .. ecc/2 + 2 errors, so some lie beyond the
    radius."""
    out = {}
    rng = np.random.default_rng(23)
    T = 8
    for n, ecc, il in IL_REPAIR_CASES:
        k = n - ecc
        ri = L.RSi16(n, ecc, il, simd_x4=False, simd_x8=False, simd_x16=False)
        msgs = rng.integers(0, 65536, size=(T, k * il), dtype=np.uint16)
        pars = np.stack([np.frombuffer(ri.encode(m.tobytes()), np.uint16) for m in msgs])
        m_in, e_in = msgs.copy(), pars.copy()
        positions = np.full((T, ecc), -1, np.int64)
        for t in range(T):
            known = t >= T // 2
            if known:
                npos = int(rng.integers(1, ecc + 1))
                pos = np.sort(rng.choice(n, size=npos, replace=False))
                positions[t, :npos] = pos
            for c in range(il):
                if known:
                    cand = positions[t][positions[t] >= 0]
                    ne = int(rng.integers(0, len(cand) + 1))
                    where = rng.choice(cand, size=ne, replace=False)
                else:
                    ne = int(rng.integers(0, ecc // 2 + 3))
                    where = rng.choice(n, size=ne, replace=False)
                for p in where:
                    v = np.uint16(rng.integers(1, 65536))
                    if p < k:
                        m_in[t, p * il + c] ^= v
                    else:
                        e_in[t, (p - k) * il + c] ^= v
        status = np.zeros(T, np.uint8)
        m_out, e_out = m_in.copy(), e_in.copy()
        for t in range(T):
            mb, eb = bytearray(m_in[t].tobytes()), bytearray(e_in[t].tobytes())
            pos = [int(p) for p in positions[t] if p >= 0]
            try:
                status[t] = int(ri.repair(mb, eb, pos) if pos else ri.repair(mb, eb))
            except RuntimeError:
                status[t] = 255
            m_out[t] = np.frombuffer(mb, np.uint16)
            e_out[t] = np.frombuffer(eb, np.uint16)
        out[f"rsi16_ilrepair_{n}_{ecc}_{il}"] = dict(msgs=m_in, eccs=e_in, positions=positions, status=status,
                                                     msgs_out=m_out, eccs_out=e_out)
    return out


def reftest_fixture(L):
    """(g) every shape of the reference's own RS256 and RSi16 tests: per shape a batch of
    messages (regenerated from a seed; digests stored), the parity CRCs (RSi16: SIMD and
    scalar words), and the status and repaired-codeword CRCs after 0 .. ecc/2 + 2 (RS256) or
    0 .. ecc/2 + 1 (RSi16) errors per codeword."""
    rs = {}
    for bl, ecc in C.RS256_REF_SHAPES:
        k = bl - ecc
        msgs, counts, s = C.rs256_reftest_case(bl, ecc)
        ref = L.RS256(block_len=bl, ecc_len=ecc)
        par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint8).reshape(-1, ecc)
        cws = np.concatenate([msgs, par], 1)
        W.rs256_corrupt(cws, counts, s)
        st, rep = ref_repair_rs256(ref, cws, k)
        t = f"{bl}_{ecc}"
        rs.update({f"in_sha_{t}": C.sha(msgs, counts), f"parity_crc_{t}": C.row_crc(par),
                   f"corrupt_sha_{t}": C.sha(cws), f"status_{t}": st, f"repaired_crc_{t}": C.row_crc(rep)})
    ri = {}
    for n, ecc in C.RSI16_REF_SHAPES:
        k = n - ecc
        msgs, counts, s = C.rsi16_reftest_case(n, ecc)
        ref = L.RSi16(n, ecc)
        par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
        sca = L.RSi16(n, ecc, simd_x4=False, simd_x8=False, simd_x16=False)
        par_s = np.frombuffer(sca.encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
        cws = np.concatenate([msgs, par], 1)
        W.rsi16_corrupt(cws, counts, s)
        st, rep = ref_repair_rsi16(ref, cws, k)
        t = f"{n}_{ecc}"
        ri.update({f"in_sha_{t}": C.sha(msgs, counts), f"parity_crc_{t}": C.row_crc(par),
                   f"parity_scalar_crc_{t}": C.row_crc(par_s), f"corrupt_sha_{t}": C.sha(cws),
                   f"status_{t}": st, f"repaired_crc_{t}": C.row_crc(rep)})
        print("rsi16 reftest", t, "statuses", np.bincount(st, minlength=5)[[0, 2, 4]], "255:", int((st == 255).sum()))
    return {"rs256_reftest": rs, "rsi16_reftest": ri}


def divergence_fixture(L):
    """(h) RSi16 shapes whose mix has a constant 65536: messages (explicit) and the parity of
    the SIMD words and of the scalar words, which differ where the message residue at that
    index is 0 mod p (and its lazy sum non-negative)."""
    d = {}
    for n, ecc in C.RSI16_DIV_SHAPES:
        msgs = C.rsi16_divergence_case(n, ecc)
        pv = np.frombuffer(L.RSi16(n, ecc).encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
        ps = np.frombuffer(L.RSi16(n, ecc, simd_x4=False, simd_x8=False, simd_x16=False)
                           .encode(msgs.tobytes()), np.uint16).reshape(-1, ecc)
        t = f"{n}_{ecc}"
        d.update({f"msgs_{t}": msgs, f"parity_{t}": pv, f"parity_scalar_{t}": ps})
        print("divergence", t, "rows whose words differ:", int((pv != ps).any(1).sum()), "of", len(msgs))
    return {"rsi16_simd_div": d}


def ilsimd_fixture(L):
    """(i) interleaved repair without error positions with the SIMD words, W = 16, 8, 4
    (simd_x16; simd_x16=False; simd_x16=simd_x8=False), where the reference repairs the columns
    in coupled groups (pyrsi16.hpp:389-453): per trial the digest of the corrupted input, the
    call's status (255: raises) and the CRCs of both buffers after the call."""
    d = {}
    flags = {16: {}, 8: dict(simd_x16=False), 4: dict(simd_x16=False, simd_x8=False)}
    for n, ecc, il in C.RSI16_ILSIMD_CASES:
        for w in C.RSI16_ILSIMD_WIDTHS:
            ref = L.RSi16(n, ecc, il, **flags[w])
            st = np.zeros(C.RSI16_ILSIMD_TRIALS, np.uint8)
            sha = np.zeros((C.RSI16_ILSIMD_TRIALS, 32), np.uint8)
            crc = np.zeros((C.RSI16_ILSIMD_TRIALS, 2), np.uint32)
            for t in range(C.RSI16_ILSIMD_TRIALS):
                msg, s = C.rsi16_ilsimd_messages(n, ecc, il, w, t)
                par = np.frombuffer(ref.encode(msg.tobytes()), np.uint16)
                m, e = C.rsi16_ilsimd_corrupt(n, ecc, il, msg, par, s, t)
                sha[t] = C.sha(m, e)
                mb, eb = bytearray(m.tobytes()), bytearray(e.tobytes())
                try:
                    st[t] = int(ref.repair(mb, eb))
                except RuntimeError:
                    st[t] = 255
                crc[t] = [C.row_crc(np.frombuffer(mb, np.uint16)[None])[0], C.row_crc(np.frombuffer(eb, np.uint16)[None])[0]]
            tag = f"{n}_{ecc}_{il}_w{w}"
            d.update({f"sha_{tag}": sha, f"status_{tag}": st, f"crc_{tag}": crc})
            # with error positions
            kst = np.zeros(C.RSI16_ILSIMD_KNOWN_TRIALS, np.uint8)
            kcrc = np.zeros((C.RSI16_ILSIMD_KNOWN_TRIALS, 2), np.uint32)
            ksha = np.zeros((C.RSI16_ILSIMD_KNOWN_TRIALS, 32), np.uint8)
            for t in range(C.RSI16_ILSIMD_KNOWN_TRIALS):
                msg, s = C.rsi16_ilsimd_messages(n, ecc, il, w, 100 + t)
                par = np.frombuffer(ref.encode(msg.tobytes()), np.uint16)
                m, e, pos = C.rsi16_ilsimd_known(n, ecc, il, msg, par, s)
                ksha[t] = C.sha(m, e, pos)
                mb, eb = bytearray(m.tobytes()), bytearray(e.tobytes())
                try:
                    kst[t] = int(ref.repair(mb, eb, [int(x) for x in pos]))
                except RuntimeError:
                    kst[t] = 255
                kcrc[t] = [C.row_crc(np.frombuffer(mb, np.uint16)[None])[0],
                           C.row_crc(np.frombuffer(eb, np.uint16)[None])[0]]
            d.update({f"ksha_{tag}": ksha, f"kstatus_{tag}": kst, f"kcrc_{tag}": kcrc})
            print("ilsimd", tag, "statuses", st.tolist())
    return {"rsi16_ilsimd": d}


def _raises(fn, *args):
    try:
        return fn(*args), 0
    except RuntimeError:
        return None, 1


def ntt_fixture(L):
    """(j) the NTT class (pyntt.hpp:28-84) and RSi16._sugiyama<W> (pyrsi16.hpp:653-674) with
    explicit inputs and the reference's outputs (u16 as the reference returns them; Sugiyama's
    u32 field elements); raises[t] = 1 where the call raises RuntimeError (outputs zero)."""
    d = {}
    s = W.Stream(47)
    for bl, ecc, prim in C.NTT_SHAPES:
        gf = L.GFi16(prim)  # the reference's NTT keeps a reference to gf: it must stay alive
        ntt = L.NTT(gf, bl, ecc)
        N, E = ntt.ntt_len, ntt.ecc_len
        for w in C.NTT_WIDTHS:
            sfx = "" if w == 1 else str(w)
            tag = f"{bl}_{ecc}_{prim}_w{w}"
            if N <= 256 or w <= 4:  # the transforms of ntt_len 1024 on 1 and 4 vectors only
                x = s.u16(w * N)
                d[f"x_{tag}"] = x
                for op in ("ntt", "intt", "nttr", "inttr"):
                    d[f"{op}_{tag}"] = np.frombuffer(getattr(ntt, op + sfx)(x.tobytes()), np.uint16)
            # poly_mul: a and b of the given length kinds
            kinds = [("half", "half"), ("full", "full"), ("rand", "rand"), ("rand", "half")]
            A, B, O = [], [], []
            for ka, kb in kinds:
                a = C.poly_vectors(s, w, E, C.poly_lengths(s, w, E, ka))
                b = C.poly_vectors(s, w, E, C.poly_lengths(s, w, E, kb))
                A.append(a), B.append(b)
                O.append(np.frombuffer(getattr(ntt, "poly_mul" + sfx)(a.tobytes(), b.tobytes()), np.uint16))
            d[f"mul_a_{tag}"], d[f"mul_b_{tag}"], d[f"mul_out_{tag}"] = np.stack(A), np.stack(B), np.stack(O)
            # poly_div: random lengths, a zero divisor, f shorter than g, b with its top two zero
            A, B, O, R = [], [], [], []
            for t in range(10):
                la = C.poly_lengths(s, w, E, "rand")
                lb = C.poly_lengths(s, w, E, "pos")
                if t == 1:
                    lb[s.ints(0, w - 1, 1)[0]] = 0          # one zero divisor
                elif t == 2:
                    la[:], lb[:] = 0, 0                      # 0 // 0
                elif t == 3:
                    la[:] = np.minimum(lb, E) - 1            # deg f < deg g
                    la[la < 0] = 0
                elif t == 4:
                    lb[:] = E - 2                            # the reference's test_poly_div_n
                    la[:] = E
                elif t == 5:
                    lb[:] = 1                                # constant divisors
                a = C.poly_vectors(s, w, E, la)
                b = C.poly_vectors(s, w, E, lb)
                o, r = _raises(getattr(ntt, "poly_div" + sfx), a.tobytes(), b.tobytes())
                A.append(a), B.append(b), R.append(r)
                O.append(np.zeros(w * E, np.uint16) if r else np.frombuffer(o, np.uint16))
            d[f"div_a_{tag}"], d[f"div_b_{tag}"] = np.stack(A), np.stack(B)
            d[f"div_out_{tag}"], d[f"div_raises_{tag}"] = np.stack(O), np.array(R, np.uint8)
            # poly_inv: every n of the list, a of half and random lengths (length >= 1)
            A, NN, O, R = [], [], [], []
            for n in C.ntt_inv_mods(E):
                for kind in ("half", "pos") if E <= 16 else ("pos",):
                    a = C.poly_vectors(s, w, E, C.poly_lengths(s, w, E, kind))
                    o, r = _raises(getattr(ntt, "poly_inv" + sfx), a.tobytes(), n)
                    A.append(a), NN.append(n), R.append(r)
                    O.append(np.zeros(w * E, np.uint16) if r else np.frombuffer(o, np.uint16))
            d[f"inv_a_{tag}"], d[f"inv_n_{tag}"] = np.stack(A), np.array(NN, np.uint32)
            d[f"inv_out_{tag}"], d[f"inv_raises_{tag}"] = np.stack(O), np.array(R, np.uint8)
            print("ntt", tag, "div raises", sum(d[f"div_raises_{tag}"]), "inv raises", sum(d[f"inv_raises_{tag}"]))

    # _sugiyama<W>: syndromes of codewords with 0 .. ecc/2 + 2 errors per vector, and random ones
    for n, ecc in C.SUGIYAMA_SHAPES:
        ref = L.RSi16(n, ecc)
        k = n - ecc
        for w in C.NTT_WIDTHS:
            tag = f"{n}_{ecc}_w{w}"
            S, LOC, EV, R = [], [], [], []
            for t in range(6):
                if t < 5:
                    msgs = s.u16(w * k).reshape(w, k)
                    par = np.frombuffer(ref.encode(msgs.tobytes()), np.uint16).reshape(w, ecc)
                    cws = np.concatenate([msgs, par], 1)
                    counts = s.ints(0, ecc // 2 + 2, w) if t else np.full(w, ecc // 2, np.int64)
                    W.rsi16_corrupt(cws, counts, s)
                    syn = np.concatenate([np.array(ref._synd(cw[:k].tobytes(), cw[k:].tobytes()), np.uint32)
                                          for cw in cws])
                else:
                    syn = s.ints(0, 65536, w * ecc).astype(np.uint32)
                fn = ref._sugiyama if w == 1 else getattr(ref, f"_sugiyama{w}")
                o, r = _raises(fn, [int(v) for v in syn])
                S.append(syn), R.append(r)
                LOC.append(np.zeros(w * ecc, np.uint32) if r else np.array(o[0], np.uint32))
                EV.append(np.zeros(w * ecc, np.uint32) if r else np.array(o[1], np.uint32))
            d[f"sug_synd_{tag}"], d[f"sug_loc_{tag}"] = np.stack(S), np.stack(LOC)
            d[f"sug_ev_{tag}"], d[f"sug_raises_{tag}"] = np.stack(EV), np.array(R, np.uint8)
            print("sugiyama", tag, "raises", R)
    return {"rsi16_ntt": d}


def circ16_fixture(L):
    """(k) CIRC16 (pycirc16.hpp:33-431) on the reference's own test shapes and flags: per case the
    encode of a seeded message (ecc explicit where small, else its CRC), the repair of each
    scenario of cases.circ16_scenarios (input digests, output CRCs, whether it raises) and the
    outer error rows of cases.circ16_find_inputs."""
    d = {}
    for i, case in enumerate(C.CIRC16_CASES):
        tag = C.circ16_tag(case)
        ref = L.CIRC16(*case[:5], **C.circ16_flags(case[5]))
        lay, msg = C.circ16_message(i, case)
        assert ref.message_len == lay.msg_len and ref.ecc_len == lay.ecc_len
        ecc = np.frombuffer(ref.encode(msg.tobytes()), np.uint16).copy()
        d[f"{tag}_msg_sha"] = C.sha(msg)
        d[f"{tag}_ecc_crc"] = C.row_crc(ecc[None])
        if ecc.size <= 8192:
            d[f"{tag}_ecc"] = ecc
        names, st = [], []
        for name, m, e in C.circ16_scenarios(i, case, lay, msg, ecc):
            mb, eb = bytearray(m.tobytes()), bytearray(e.tobytes())
            d[f"{tag}_{name}_in_sha"] = C.sha(m, e)
            try:
                ref.repair(mb, eb)
                raised = 0
            except RuntimeError:
                raised = 1
            d[f"{tag}_{name}_raised"] = np.array([raised], np.uint8)
            d[f"{tag}_{name}_out_crc"] = C.row_crc(np.frombuffer(bytes(mb), np.uint16)[None]).tolist() + \
                C.row_crc(np.frombuffer(bytes(eb), np.uint16)[None]).tolist()
            d[f"{tag}_{name}_out_crc"] = np.array(d[f"{tag}_{name}_out_crc"], np.uint32)
            names.append(name)
            st.append(raised)
        finds = []
        for il, m in C.circ16_find_inputs(i, case, lay, msg, ecc):
            rows = ref._find_outer_error_locations(bytearray(m.tobytes()), bytearray(ecc.tobytes()), il)
            finds.append(list(rows))
        d[f"{tag}_find"] = np.array([r + [-1] * (lay.on + 1 - len(r)) for r in finds], np.int32)
        print("circ16", tag, dict(zip(names, st)), [len(f) for f in finds])
    return {"circ16": d}


def _generator_rows(L, n, ecc):
    """G (k x ecc, canonical) of RSi16(n, ecc) in the scalar words: row m = the parity of the
    unit message at m (a 0 that is really 65536 recovered from the parity of 2 e_m)."""
    rs = L.RSi16(n, ecc, simd_x4=False, simd_x8=False, simd_x16=False)
    k = n - ecc
    eye = np.eye(k, dtype=np.uint16)
    g1 = np.frombuffer(rs.encode(eye.tobytes()), np.uint16).reshape(k, ecc).astype(np.int64)
    g2 = np.frombuffer(rs.encode((2 * eye).astype(np.uint16).tobytes()), np.uint16).reshape(k, ecc).astype(np.int64)
    return np.where((g1 == 0) & (g2 == 65535), 65536, g1)


def _parity65536(G, msgs):
    """rows of msgs (u16) whose linear parity holds a 65536 somewhere"""
    v = msgs.astype(np.int64)
    par = np.zeros((len(msgs), G.shape[1]), np.int64)
    for j in range(0, G.shape[0], 64):  # exact in int64: 64 * 65535 * 65536 < 2^63
        par = (par + v[:, j:j + 64] @ G[j:j + 64]) % 65537
    return np.nonzero((par == 65536).any(1))[0]


def circ16_planted_fixture(L):
    """(l) CIRC16 blocks whose outer parity holds a 65536 (stored 0, its inner parity encodes
    65536: the repair's outer-zero step) and blocks whose inner parity holds one (the inner-zero
    step), found with the linear map and confirmed by the reference; explicit messages, the
    reference's ecc, and the repairs of the clean block and of corrupted rows."""
    d = {}
    for i, case in enumerate(C.CIRC16_PLANT_CASES):
        tag = C.circ16_tag(case)
        ibl, iecc, obl, oecc, il, w = case
        ref = L.CIRC16(*case[:5], **C.circ16_flags(w))
        lay = C.CircLayout(*case[:5])
        s = W.Stream(1900 + i)
        msgs = {}
        # outer: a column whose outer parity holds 65536
        Go = _generator_rows(L, obl, oecc)
        for _ in range(64):
            cols = s.u16(65536 * lay.ok).reshape(65536, lay.ok)
            hit = _parity65536(Go, cols)
            if hit.size:
                break
        c0 = int(s.ints(0, lay.rI - 1, 1)[0])
        msgs["outer"] = C.circ16_planted_message(i, "outer", lay, c0, cols[hit[0]])
        d[f"{tag}_outer_plant"] = np.concatenate([[c0], cols[hit[0]]]).astype(np.uint32)
        # inner: a message row whose inner parity holds 65536
        Gi = _generator_rows(L, ibl, iecc)
        for _ in range(64):
            rows = s.u16(65536 * lay.ik).reshape(65536, lay.ik)
            hit = _parity65536(Gi, rows)
            if hit.size:
                break
        q = int(s.ints(0, lay.ok * lay.I - 1, 1)[0])
        msgs["inner"] = C.circ16_planted_message(i, "inner", lay, q, rows[hit[0]])
        d[f"{tag}_inner_plant"] = np.concatenate([[q], rows[hit[0]]]).astype(np.uint32)
        for kind, m in msgs.items():
            ecc = np.frombuffer(ref.encode(m.tobytes()), np.uint16).copy()
            pre = f"{tag}_{kind}"
            d[f"{pre}_msg_sha"], d[f"{pre}_ecc"] = C.sha(m), ecc
            if kind == "outer":  # confirmed: the inner parity of the outer parity rows differs
                rso = ecc[:lay.rso_len]
                rsio = ecc[lay.ecc_len - lay.iE * lay.oE * lay.I:]
                assert bytes(ref.rsi.encode(rso.tobytes())) != rsio.tobytes(), tag
            else:
                k0 = q * lay.ik
                e0 = lay.rso_len + q * lay.iE
                sy = ref.rsi._synd(m[k0:k0 + lay.ik].tobytes(), ecc[e0:e0 + lay.iE].tobytes())
                assert any(sy), tag
            restored = {}
            for name, mm, ee in C.circ16_planted_scenarios(i, kind, lay, m, ecc):
                mb, eb = bytearray(mm.tobytes()), bytearray(ee.tobytes())
                d[f"{pre}_{name}_in_sha"] = C.sha(mm, ee)
                try:
                    ref.repair(mb, eb)
                    raised = 0
                except RuntimeError:
                    raised = 1
                d[f"{pre}_{name}_raised"] = np.array([raised], np.uint8)
                d[f"{pre}_{name}_out_crc"] = np.array(
                    [C.row_crc(np.frombuffer(bytes(mb), np.uint16)[None])[0],
                     C.row_crc(np.frombuffer(bytes(eb), np.uint16)[None])[0]], np.uint32)
                restored[name] = bytes(mb) == m.tobytes() and bytes(eb) == ecc.tobytes()
            rows = ref._find_outer_error_locations(bytearray(m.tobytes()), bytearray(ecc.tobytes()), 0)
            d[f"{pre}_find0"] = np.array(list(rows) + [-1], np.int32)
            print("circ16 planted", pre, "find0", list(rows), "restored", restored)
    return {"circ16_planted": d}


if __name__ == "__main__":
    if sys.argv[1:] == ["interleave"]:
        for key, v in interleave_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["interleave_repair"]:
        for key, v in interleave_repair_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["ilsimd"]:
        for key, v in ilsimd_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["reftest"]:
        for key, v in reftest_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["divergence"]:
        for key, v in divergence_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["circ16"]:
        for key, v in circ16_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["circ16_planted"]:
        for key, v in circ16_planted_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["ntt"]:
        for key, v in ntt_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["radius1024"]:
        for key, v in radius1024_fixture(load_reference()).items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    elif sys.argv[1:] == ["radius"]:
        from libffrs_amd import workloads as W  # noqa: F811
        d = radius_fixture(load_reference())
        for key, v in d.items():
            np.savez_compressed(os.path.join(HERE, key + ".npz"), **v)
    else:
        main()
