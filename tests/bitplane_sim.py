"""NumPy evaluator of an encoded bit-plane program (test utility).

Mirrors what the runtime's bit-plane kernel does with the blob -- input
planes from the stimulus generator, levelised gates over 64-lane planes,
unmasked commit, synchronous memory ports at the edge (per-lane reads of the
old words, then guarded writes), per-lane output hashing, write-back -- so
the lowering, scheduling and encoding can be checked on the CPU against the
oracle."""
import numpy as np

from rtlflow_amd import bitplane as bp

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix(x):
    with np.errstate(over="ignore"):
        z = x + GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def stim(seed, sids, c, k):
    with np.errstate(over="ignore"):
        x = np.uint64(seed) ^ (sids.astype(np.uint64) * GOLD) ^ np.uint64((c << 8) & 0xFFFFFFFFFFFFFFFF) \
            ^ np.uint64(k)
    return splitmix(x)


def to_planes(vals, nbits):
    """vals: uint64[64] -> planes[nbits] (bit t of plane b = bit b of vals[t])."""
    bits = (vals[None, :] >> np.arange(nbits, dtype=np.uint64)[:, None]) & np.uint64(1)
    return (bits << np.arange(64, dtype=np.uint64)[None, :]).sum(axis=1, dtype=np.uint64)


def from_planes(planes):
    """planes[nbits] -> uint64[64] lane values (nbits <= 64)."""
    nb = len(planes)
    bits = (planes[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)
    return (bits << np.arange(nb, dtype=np.uint64)[:, None]).sum(axis=0, dtype=np.uint64)


def fnv_bytes(h, v, nbytes):
    with np.errstate(over="ignore"):
        for i in range(nbytes):
            h = (h ^ ((v >> np.uint64(8 * i)) & np.uint64(0xFF))) * np.uint64(0x100000001B3)
    return h


class Program:
    def __init__(self, blob):
        self.b = np.asarray(blob, dtype=np.uint64)
        self.sec = bp.sections(blob)
        h = blob[:bp.HDR_WORDS]
        self.nslots, self.n_in, self.n_state, self.n_commit, self.n_hash = h[2:7]
        self.nlev = h[8:11]
        self.ngates = h[11:14]
        self.byte_form = bool(h[1] >> 16)
        b = self.b
        s = self.sec
        self.inputs = b[s["in"]:s["in"] + 4 * self.n_in].reshape(-1, 4).astype(np.int64)
        self.state = b[s["state"]:s["state"] + 6 * self.n_state].reshape(-1, 6).astype(np.int64)
        ncm = (self.n_commit, blob[16])
        self.commits = [b[s[nm]:s[nm] + 2 * ncm[k]].reshape(-1, 2).astype(np.int64)
                        for k, nm in enumerate(("commitA0", "commitM"))]
        self.phases = [self._phase(s["lev" + nm], self.nlev[k], s["gates" + nm], self.ngates[k])
                       for k, nm in enumerate(("A0", "M", "B"))]
        self.hashes = []
        for nm in ("hashM", "hashB"):
            o = s[nm]
            hp = []
            for _ in range(self.n_hash):
                nbytes, width = int(b[o]), int(b[o + 1])
                hp.append((nbytes, width, b[o + 2:o + 2 + width].astype(np.int64)))
                o += 2 + width
            self.hashes.append(hp)
        self.mems = [tuple(int(x) for x in b[s["mems"] + 8 * i:s["mems"] + 8 * i + 8])
                     for i in range(blob[17])]
        self.ports = []
        for nm in ("portsA0", "portsM"):
            o = s[nm]
            nrec = int(b[o + 1])
            o += 2
            recs = []
            for _ in range(nrec):
                kind, mi, na, n2 = (int(x) for x in b[o:o + 4])
                if kind == 0:
                    addr = b[o + 4:o + 4 + na].astype(np.int64)
                    tg = [(int(x) >> 24, int(x) & 0xFFFFFF) for x in b[o + 4 + na:o + 4 + na + n2]]
                    recs.append((0, mi, addr, tg))
                    o += 4 + na + n2
                else:
                    g = int(b[o + 4])
                    addr = b[o + 5:o + 5 + na].astype(np.int64)
                    data = b[o + 5 + na:o + 5 + na + n2].astype(np.int64)
                    recs.append((1, mi, addr, g, data))
                    o += 5 + na + n2
            self.ports.append(recs)
        self.fchunks = []
        for c in s["fchunks"]:
            ph = self._phase(c["lev"], c["nlev"], c["gates"], c["ngates"])
            o = c["wb"]
            wb = []
            for _ in range(c["n_wb"]):
                cls, lo, hi, words, width = (int(x) for x in b[o:o + 5])
                wb.append((cls, lo | (hi << 32), words, width, b[o + 5:o + 5 + width].astype(np.int64)))
                o += 5 + width
            self.fchunks.append((ph, wb))

    def _phase(self, lev_at, nlev, gates_at, ngates):
        b = self.b
        offs = b[lev_at:lev_at + nlev + 1].astype(np.int64)
        g = b[gates_at:gates_at + 2 * ngates].reshape(-1, 2)
        w = g[:, 0] | (g[:, 1] << np.uint64(32))
        f = lambda sh, bits: ((w >> np.uint64(sh)) & np.uint64((1 << bits) - 1)).astype(np.int64)
        if self.byte_form:
            # LDS byte offsets, inversion flags in bit 0 of the x and y fields
            return offs, (f(0, 16) >> 3, f(16, 16) >> 3, f(32, 16) >> 3, f(32, 1), f(48, 16) >> 3,
                          f(48, 1))
        return offs, (f(0, 16), f(16, 16), f(32, 15), f(47, 1), f(48, 15), f(63, 1))

    def run_phase(self, L, k):
        offs, (dst, s, xs, ix, ys, iy) = self.phases[k] if isinstance(k, int) else k
        for l in range(len(offs) - 1):
            a, e = offs[l], offs[l + 1]
            S = L[s[a:e]]
            X = L[xs[a:e]] ^ (M64 * ix[a:e].astype(np.uint64))
            Y = L[ys[a:e]] ^ (M64 * iy[a:e].astype(np.uint64))
            L[dst[a:e]] = (S & X) | (~S & Y)

    def ref(self, L, r):
        v = L[r & 0x7FFFFFFF]
        return v ^ M64 if r >> 31 else v

    def inputs_to(self, L, seed, sids, c):
        for k, nb, sa, sb in self.inputs:
            L[sa:sa + nb] = to_planes(stim(seed, sids, c, int(k)), int(nb))

    def copy_inputs(self, L):
        for k, nb, sa, sb in self.inputs:
            L[sb:sb + nb] = L[sa:sa + nb]

    def commit(self, L, k, act):
        # unmasked: lanes that are not active never write their state back
        st, dn = self.commits[k][:, 0], self.commits[k][:, 1]
        L[st] = L[dn]

    def new_mems(self):
        """Zeroed memories of one 64-lane group: [64, depth] per memory."""
        return [np.zeros((64, m[6]), np.uint64) for m in self.mems]

    def lane_vals(self, L, refs):
        return from_planes(np.array([self.ref(L, int(r)) for r in refs], np.uint64))

    def ram_step(self, L, k, act, mem):
        lanes = np.arange(64)
        on = ((np.uint64(act) >> lanes.astype(np.uint64)) & np.uint64(1)) == 1
        for rec in self.ports[k]:
            mi = rec[1]
            depth = self.mems[mi][6]
            m = mem[mi]
            addr = self.lane_vals(L, rec[2]).astype(np.int64)
            ok = addr < depth
            if rec[0] == 0:
                old = np.where(ok, m[lanes, np.where(ok, addr, 0)], np.uint64(0))
                planes = to_planes(old, 64)
                for bit, slot in rec[3]:
                    L[slot] = planes[bit]
            else:
                g = ((self.ref(L, rec[3]) >> lanes.astype(np.uint64)) & np.uint64(1)) == 1
                data = self.lane_vals(L, rec[4])
                w = g & on & ok
                m[lanes[w], addr[w]] = data[w]

    def hash(self, L, k, h):
        for nbytes, width, refs in self.hashes[k]:
            planes = np.array([self.ref(L, int(r)) for r in refs], np.uint64)
            if width <= 64:
                h = fnv_bytes(h, from_planes(planes), nbytes)
            else:
                for w0 in range(0, nbytes * 8, 32):
                    h = fnv_bytes(h, from_planes(planes[w0:w0 + 32]), 4)
        return h

    def cycles(self, L, seed, sids, ncycles, act, h, mem=None):
        """The kernel's schedule: A0, commit and memory ports; then per cycle
        the merged pass M (outputs of cycle n, next state of cycle n + 1),
        commit and memory ports, hash of cycle n; B and the hash of the last
        cycle."""
        if ncycles == 0:
            return h
        if mem is None:
            mem = self.new_mems()
        self.inputs_to(L, seed, sids, 0)
        self.run_phase(L, 0)
        self.commit(L, 0, act)
        self.ram_step(L, 0, act, mem)
        self.copy_inputs(L)
        if ncycles > 1:
            self.inputs_to(L, seed, sids, 1)
        for n in range(ncycles - 1):
            self.run_phase(L, 1)
            self.commit(L, 1, act)
            self.ram_step(L, 1, act, mem)
            h = self.hash(L, 0, h)
            self.copy_inputs(L)
            if n + 2 < ncycles:
                self.inputs_to(L, seed, sids, n + 2)
        self.run_phase(L, 2)
        return self.hash(L, 1, h)

    def run(self, seed, nlanes, ncycles, stim0=0):
        """Lanes start from zero state and run ``ncycles`` cycles; returns
        hashes[nlanes]."""
        out = np.zeros(nlanes, np.uint64)
        for g0 in range(0, nlanes, 64):
            n = min(64, nlanes - g0)
            act = np.uint64((1 << n) - 1) if n < 64 else M64
            sids = np.arange(g0 + stim0, g0 + stim0 + 64, dtype=np.uint64)
            L = np.zeros(self.nslots, np.uint64)
            h = self.cycles(L, seed, sids, ncycles, act, np.full(64, 0xCBF29CE484222325, np.uint64))
            out[g0:g0 + n] = h[:n]
        return out


def run_state(prog: "Program", seed, nlanes, ncycles, lane):
    """Run like ``Program.run`` and return the write-back values of one lane
    as {(cls, loc): value}."""
    g0 = (lane // 64) * 64
    n = min(64, nlanes - g0)
    act = np.uint64((1 << n) - 1) if n < 64 else M64
    sids = np.arange(g0, g0 + 64, dtype=np.uint64)
    L = np.zeros(prog.nslots, np.uint64)
    prog.cycles(L, seed, sids, ncycles, act, np.full(64, 0xCBF29CE484222325, np.uint64))
    t = lane - g0
    out = {}
    for ph, wb in prog.fchunks:
        prog.run_phase(L, ph)
        for cls, loc, words, width, refs in wb:
            bits = [int((prog.ref(L, int(r)) >> np.uint64(t)) & np.uint64(1)) for r in refs]
            out[(cls, loc)] = sum(b << i for i, b in enumerate(bits))
    return out
