"""Codec backends for the golden replay (tests/golden/replay.py).

OracleBackend drives the CPU restatement (oracle/); GpuBackend drives the HIP
path through the product's public API (libffrs_amd), the same calls a user makes.
"""
import numpy as np

from oracle import oracle as O


class OracleBackend:
    def __init__(self):
        self._rs = {}
        self._ri = {}

    def rs(self, ecc):
        if ecc not in self._rs:
            self._rs[ecc] = O.RS256(ecc)
        return self._rs[ecc]

    def ri(self, n, ecc):
        if (n, ecc) not in self._ri:
            self._ri[(n, ecc)] = O.RSi16(n, ecc)
        return self._ri[(n, ecc)]

    def rs256_encode(self, block_len, ecc, buf):
        return self.rs(ecc).encode(buf, block_len - ecc)

    def rs256_encode_batch(self, ecc, msgs):
        return self.rs(ecc).encode_batch(msgs)

    def rs256_repair_batch(self, ecc, msgs, parity):
        return self.rs(ecc).decode_batch(msgs, parity)

    def rs256_synds(self, block_len, ecc, cw):
        k = block_len - ecc
        return self.rs(ecc).synds(cw[:k], cw[k:])

    def rs256_mixed_encode(self, eccs, lens, msgs):
        return [self.rs(int(e)).encode_batch(m[None])[0] for e, m in zip(eccs, msgs)]

    def rs256_mixed_repair(self, eccs, lens, cws):
        st = np.zeros(len(cws), np.uint8)
        for i, (e, l) in enumerate(zip(eccs, lens)):
            m = cws[i][:l].copy()
            p = cws[i][l:].copy()
            st[i] = self.rs(int(e)).decode(m, p)
            cws[i][:l] = m
            cws[i][l:] = p
        return st

    def rsi16_encode_batch(self, n, ecc, msgs, vector):
        return self.ri(n, ecc).encode(msgs, vector)

    def rsi16_repair_batch(self, n, ecc, msgs, parity):
        return self.ri(n, ecc).repair_batch(msgs, parity)

    def rsi16_synds(self, n, ecc, msg, parity):
        return self.ri(n, ecc).synds(msg, parity)

    def rsi16_encode_interleaved(self, n, ecc, il, blocks, vector=True):
        return _interleaved_by_columns(lambda c: self.ri(n, ecc).encode(c, vector), n, ecc, il, blocks)

    def rsi16_repair_interleaved(self, n, ecc, il, msg, parity, positions=None):
        """The scalar-word interleaved repair as the reference runs it (pyrsi16.hpp:391-453 with
        SIMD_W = 1): column after column, each written back after its repair, the status the
        maximum; a column whose repair raises ends the call (255) with the columns before it
        written back."""
        k = n - ecc
        M = msg.reshape(k, il)
        P = parity.reshape(ecc, il)
        status = 0
        for c in range(il):
            m = np.ascontiguousarray(M[:, c])
            p = np.ascontiguousarray(P[:, c])
            st = (self.ri(n, ecc).repair_known(m, p, positions) if positions is not None
                  else self.ri(n, ecc).repair(m, p))
            if st == 255:
                return 255
            M[:, c] = m
            P[:, c] = p
            status = max(status, st)
        return status

    def rsi16_repair_interleaved_simd(self, n, ecc, il, width, msg, parity, positions=None):
        if positions is not None:  # the vector words give every column its own repair here
            return self.rsi16_repair_interleaved(n, ecc, il, msg, parity, positions)
        return self.ri(n, ecc).repair_interleaved_simd(width, msg, parity, il)

    def rsi16_sugiyama(self, n, ecc, synds):
        L, loc, ev = self.ri(n, ecc).sugiyama(synds)
        return None if L == 0xFFFFFFFF else (loc, ev)

    def rsi16_repair_known(self, n, ecc, msg, parity, positions):
        return self.ri(n, ecc).repair_known(msg, parity, positions)

    def rsi16_mixed_encode(self, ns, eccs, msgs):
        return [self.ri(int(n), int(e)).encode(m[None], True)[0] for n, e, m in zip(ns, eccs, msgs)]

    def rsi16_mixed_repair(self, ns, eccs, cws):
        st = np.zeros(len(cws), np.uint8)
        for i, (n, e) in enumerate(zip(ns, eccs)):
            k = int(n - e)
            m = cws[i][:k].copy()
            p = cws[i][k:].copy()
            st[i] = self.ri(int(n), int(e)).repair(m, p)
            cws[i][:k] = m
            cws[i][k:] = p
        return st


def _interleaved_by_columns(encode, n, ecc, il, blocks):
    """Interleaved encode through a per-codeword encoder: column c of a block is symbols
    c, c + il, c + 2 il, ... of it; parity symbol j of column c goes to j * il + c."""
    k = n - ecc
    blocks = np.asarray(blocks, np.uint16).reshape(-1, k, il)
    cols = np.ascontiguousarray(blocks.transpose(0, 2, 1).reshape(-1, k))
    par = encode(cols).reshape(-1, il, ecc)
    return np.ascontiguousarray(par.transpose(0, 2, 1)).reshape(-1, ecc * il)


class GpuBackend:
    """The HIP path through libffrs_amd's public API."""

    def __init__(self, repair_mode=0):
        import libffrs_amd
        self.L = libffrs_amd
        self._rs = {}
        self._ri = {}
        self.repair_mode = repair_mode

    def rs(self, ecc, block_len=255):
        key = ecc
        if key not in self._rs:
            self._rs[key] = self.L.RS256(block_len=255 if ecc < 255 else None, ecc_len=ecc)
            self._rs[key]._set_repair_mode(self.repair_mode)
        r = self._rs[key]
        r.block_len = block_len
        return r

    def ri(self, n, ecc, vector=True):
        key = (n, ecc, vector)
        if key not in self._ri:
            kw = {} if vector else dict(simd_x4=False, simd_x8=False, simd_x16=False)
            self._ri[key] = self.L.RSi16(n, ecc, **kw)
        return self._ri[key]

    def rs256_encode(self, block_len, ecc, buf):
        return self.rs(ecc, block_len).encode(buf)

    def rs256_encode_batch(self, ecc, msgs):
        count, k = msgs.shape
        out = self.rs(ecc, k + ecc).encode(np.ascontiguousarray(msgs).tobytes())
        return np.frombuffer(bytes(out), np.uint8).reshape(count, ecc).copy()

    def rs256_repair_batch(self, ecc, msgs, parity):
        count, k = msgs.shape
        st = self.rs(ecc, k + ecc).repair_batch(msgs, parity)
        return np.frombuffer(bytes(st), np.uint8).copy()

    def rs256_synds(self, block_len, ecc, cw):
        return np.frombuffer(bytes(self.rs(ecc, block_len)._synds(np.ascontiguousarray(cw).tobytes())), np.uint8)

    def rs256_mixed_encode(self, eccs, lens, msgs):
        out = [None] * len(msgs)
        groups = {}
        for i, (e, l) in enumerate(zip(eccs, lens)):
            groups.setdefault((int(e), int(l)), []).append(i)
        for (e, l), idx in groups.items():
            par = self.rs256_encode_batch(e, np.stack([msgs[i] for i in idx]))
            for j, i in enumerate(idx):
                out[i] = par[j]
        return out

    def rs256_mixed_repair(self, eccs, lens, cws):
        st = np.zeros(len(cws), np.uint8)
        groups = {}
        for i, (e, l) in enumerate(zip(eccs, lens)):
            groups.setdefault((int(e), int(l)), []).append(i)
        for (e, l), idx in groups.items():
            m = np.ascontiguousarray(np.stack([cws[i][:l] for i in idx]))
            p = np.ascontiguousarray(np.stack([cws[i][l:] for i in idx]))
            s = self.rs256_repair_batch(e, m, p)
            for j, i in enumerate(idx):
                st[i] = s[j]
                cws[i][:l] = m[j]
                cws[i][l:] = p[j]
        return st

    def rsi16_encode_batch(self, n, ecc, msgs, vector):
        msgs = np.ascontiguousarray(msgs, np.uint16)
        out = self.ri(n, ecc, vector).encode(msgs.tobytes())
        return np.frombuffer(bytes(out), np.uint16).reshape(-1, ecc).copy()

    def rsi16_repair_batch(self, n, ecc, msgs, parity):
        # the reference sizes buffers in bytes (util.hpp:30-55), so u16 arrays go in as byte views
        st = self.ri(n, ecc).repair_batch(msgs.view(np.uint8), parity.view(np.uint8))
        return np.frombuffer(bytes(st), np.uint8).copy()

    def rsi16_encode_interleaved(self, n, ecc, il, blocks, vector=True):
        kw = {} if vector else dict(simd_x4=False, simd_x8=False, simd_x16=False)
        key = (n, ecc, il, vector)
        if key not in self._ri:
            self._ri[key] = self.L.RSi16(n, ecc, il, **kw)
        out = self._ri[key].encode(np.ascontiguousarray(blocks, np.uint16).tobytes())
        return np.frombuffer(bytes(out), np.uint16).reshape(-1, ecc * il).copy()

    def rsi16_repair_interleaved(self, n, ecc, il, msg, parity, positions=None):
        key = (n, ecc, il, False)
        if key not in self._ri:
            self._ri[key] = self.L.RSi16(n, ecc, il, simd_x4=False, simd_x8=False, simd_x16=False)
        kw = {} if positions is None else dict(error_pos=[int(v) for v in positions])
        try:
            return int(self._ri[key].repair(msg.view(np.uint8), parity.view(np.uint8), **kw))
        except RuntimeError:
            return 255

    def rsi16_repair_interleaved_simd(self, n, ecc, il, width, msg, parity, positions=None):
        """RSi16(n, ecc, il) with the SIMD words of `width` lanes (flags set explicitly, so the
        result does not depend on the host CPU)."""
        flags = {16: dict(simd_x16=True), 8: dict(simd_x16=False, simd_x8=True),
                 4: dict(simd_x16=False, simd_x8=False, simd_x4=True)}[width]
        key = (n, ecc, il, "simd", width)
        if key not in self._ri:
            self._ri[key] = self.L.RSi16(n, ecc, il, **flags)
        kw = {} if positions is None else dict(error_pos=[int(v) for v in positions])
        try:
            return int(self._ri[key].repair(msg.view(np.uint8), parity.view(np.uint8), **kw))
        except RuntimeError:
            return 255

    def rsi16_synds(self, n, ecc, msg, parity):
        return np.array(self.ri(n, ecc)._synd(np.ascontiguousarray(msg, np.uint16).tobytes(),
                                              np.ascontiguousarray(parity, np.uint16).tobytes()), np.uint32)

    def rsi16_sugiyama(self, n, ecc, synds):
        try:
            loc, ev = self.ri(n, ecc)._sugiyama([int(v) for v in synds])
        except RuntimeError:
            return None
        return np.array(loc, np.uint32), np.array(ev, np.uint32)

    def rsi16_repair_known(self, n, ecc, msg, parity, positions):
        try:
            return int(self.ri(n, ecc).repair(msg.view(np.uint8), parity.view(np.uint8), error_pos=list(positions)))
        except RuntimeError:
            return 255

    def rsi16_mixed_encode(self, ns, eccs, msgs):
        out = [None] * len(msgs)
        groups = {}
        for i, (n, e) in enumerate(zip(ns, eccs)):
            groups.setdefault((int(n), int(e)), []).append(i)
        for (n, e), idx in groups.items():
            par = self.rsi16_encode_batch(n, e, np.stack([msgs[i] for i in idx]), True)
            for j, i in enumerate(idx):
                out[i] = par[j]
        return out

    def rsi16_mixed_repair(self, ns, eccs, cws):
        st = np.zeros(len(cws), np.uint8)
        groups = {}
        for i, (n, e) in enumerate(zip(ns, eccs)):
            groups.setdefault((int(n), int(e)), []).append(i)
        for (n, e), idx in groups.items():
            k = n - e
            m = np.ascontiguousarray(np.stack([cws[i][:k] for i in idx]))
            p = np.ascontiguousarray(np.stack([cws[i][k:] for i in idx]))
            s = self.rsi16_repair_batch(n, e, m, p)
            for j, i in enumerate(idx):
                st[i] = s[j]
                cws[i][:k] = m[j]
                cws[i][k:] = p[j]
        return st


class PlanBackend(GpuBackend):
    """Mixed batches through the plan entry points: one persistent launch per field over a
    work list, codewords stored in a shuffled order with gaps between them. Repair runs
    eagerly and then three times from a captured HIP graph (inputs restored in between);
    every replay must equal the eager result."""

    def __init__(self, graph_replays=3, shape=None):
        super().__init__()
        from libffrs_amd import capi
        self.capi = capi
        self.ctx = capi.Context(0)
        self.replays = graph_replays
        self.shape = shape  # called with every new plan: per-class kernel shapes

    def _shaped(self, plan):
        if self.shape:
            self.shape(plan)
        return plan

    def _layout(self, sizes, seed):
        rng = np.random.default_rng(seed)
        order = rng.permutation(len(sizes))
        off = np.zeros(len(sizes), np.uint64)
        pos = 0
        for i in order:
            pos += int(rng.integers(0, 3))  # small gaps, odd alignments
            off[i] = pos
            pos += int(sizes[i])
        return off, pos + 8

    def _dev(self, arr):
        b = self.ctx.alloc(max(arr.nbytes, 1))
        if arr.nbytes:
            b.upload(np.ascontiguousarray(arr))
        return b

    def _run_mixed(self, make_plan, elem, dtype, k_of, e_of, msgs_list, pars_list=None, repair=False):
        n_items = len(msgs_list)
        moff, mtot = self._layout([k_of(i) for i in range(n_items)], 1)
        eoff, etot = self._layout([e_of(i) for i in range(n_items)], 2)
        hm = np.zeros(mtot, dtype)
        he = np.zeros(etot, dtype)
        for i in range(n_items):
            hm[int(moff[i]):int(moff[i]) + k_of(i)] = msgs_list[i]
            if pars_list is not None:
                he[int(eoff[i]):int(eoff[i]) + e_of(i)] = pars_list[i]
        results = []
        for graph in (False, True):
            plan = make_plan(moff, eoff, graph)
            for rep in range(1 if not graph else self.replays):
                dm, de, ds = self._dev(hm), self._dev(he), self.ctx.alloc(n_items)
                if repair:
                    plan.repair(dm, de, ds)
                else:
                    plan.encode(dm, de)
                self.ctx.sync()
                om, oe, st = np.empty_like(hm), np.empty_like(he), np.zeros(n_items, np.uint8)
                dm.download(om)
                de.download(oe)
                ds.download(st)
                for b in (dm, de, ds):
                    b.free()
                results.append((om, oe, st))
            plan.close()
        for r in results[1:]:
            assert (r[0] == results[0][0]).all() and (r[1] == results[0][1]).all()
            if repair:
                assert (r[2] == results[0][2]).all()
        om, oe, st = results[0]
        outs_m = [om[int(moff[i]):int(moff[i]) + k_of(i)] for i in range(n_items)]
        outs_e = [oe[int(eoff[i]):int(eoff[i]) + e_of(i)] for i in range(n_items)]
        return outs_m, outs_e, st

    def rs256_mixed_encode(self, eccs, lens, msgs):
        mk = lambda mo, eo, g: self._shaped(self.capi.RS256Plan(self.ctx, eccs, lens, mo, eo, graph=g))  # noqa: E731
        _, pars, _ = self._run_mixed(mk, 1, np.uint8, lambda i: int(lens[i]), lambda i: int(eccs[i]), msgs)
        return pars

    def rs256_mixed_repair(self, eccs, lens, cws):
        mk = lambda mo, eo, g: self._shaped(self.capi.RS256Plan(self.ctx, eccs, lens, mo, eo, graph=g))  # noqa: E731
        ms = [c[:int(l)] for c, l in zip(cws, lens)]
        ps = [c[int(l):] for c, l in zip(cws, lens)]
        om, oe, st = self._run_mixed(mk, 1, np.uint8, lambda i: int(lens[i]), lambda i: int(eccs[i]), ms, ps,
                                     repair=True)
        for i, c in enumerate(cws):
            c[:int(lens[i])] = om[i]
            c[int(lens[i]):] = oe[i]
        return st

    def rsi16_mixed_encode(self, ns, eccs, msgs):
        mk = lambda mo, eo, g: self._shaped(self.capi.RSi16Plan(self.ctx, ns, eccs, mo, eo, graph=g))  # noqa: E731
        _, pars, _ = self._run_mixed(mk, 2, np.uint16, lambda i: int(ns[i] - eccs[i]), lambda i: int(eccs[i]), msgs)
        return pars

    def rsi16_mixed_repair(self, ns, eccs, cws):
        mk = lambda mo, eo, g: self._shaped(self.capi.RSi16Plan(self.ctx, ns, eccs, mo, eo, graph=g))  # noqa: E731
        ks = [int(n - e) for n, e in zip(ns, eccs)]
        ms = [c[:k] for c, k in zip(cws, ks)]
        ps = [c[k:] for c, k in zip(cws, ks)]
        om, oe, st = self._run_mixed(mk, 2, np.uint16, lambda i: ks[i], lambda i: int(eccs[i]), ms, ps, repair=True)
        for i, c in enumerate(cws):
            c[:ks[i]] = om[i]
            c[ks[i]:] = oe[i]
        return st

    # uniform batches through the same plans (one class), so that every fixture set
    # also exercises the work-queue kernels, including the variant for ecc_len > 64
    def rs256_encode_batch(self, ecc, msgs):
        count, k = msgs.shape
        par = self.rs256_mixed_encode(np.full(count, ecc), np.full(count, k), list(msgs))
        return np.array(par, np.uint8).reshape(count, ecc)

    def rs256_repair_batch(self, ecc, msgs, parity):
        count, k = msgs.shape
        cws = [np.concatenate([msgs[i], parity[i]]) for i in range(count)]
        st = self.rs256_mixed_repair(np.full(count, ecc), np.full(count, k), cws)
        for i in range(count):
            msgs[i] = cws[i][:k]
            parity[i] = cws[i][k:]
        return st

    def rsi16_encode_batch(self, n, ecc, msgs, vector=True):
        if not vector:
            return super().rsi16_encode_batch(n, ecc, msgs, vector)
        count = msgs.shape[0]
        par = self.rsi16_mixed_encode(np.full(count, n), np.full(count, ecc), list(msgs))
        return np.array(par, np.uint16).reshape(count, ecc)

    def rsi16_repair_batch(self, n, ecc, msgs, parity):
        count, k = msgs.shape
        cws = [np.concatenate([msgs[i], parity[i]]) for i in range(count)]
        st = self.rsi16_mixed_repair(np.full(count, n), np.full(count, ecc), cws)
        for i in range(count):
            msgs[i] = cws[i][:k]
            parity[i] = cws[i][k:]
        return st
