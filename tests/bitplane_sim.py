"""NumPy evaluator of an encoded bit-plane program (test utility).

Mirrors what the runtime's bit-plane kernel does with the blob -- input
planes from the stimulus generator, levelised gates over 64-lane planes,
masked commit, per-lane output hashing, write-back -- so the lowering,
scheduling and encoding can be checked on the CPU against the oracle."""
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
        self.nlev = h[8:10]
        self.ngates = h[10:12]
        b = self.b
        s = self.sec
        self.inputs = b[s["in"]:s["in"] + 3 * self.n_in].reshape(-1, 3).astype(np.int64)
        self.state = b[s["state"]:s["state"] + 6 * self.n_state].reshape(-1, 6).astype(np.int64)
        self.commit = b[s["commit"]:s["commit"] + 2 * self.n_commit].reshape(-1, 2).astype(np.int64)
        self.phases = [self._phase(s["lev" + nm], self.nlev[k], s["gates" + nm], self.ngates[k])
                       for k, nm in enumerate("AB")]
        o = s["hash"]
        self.hash = []
        for _ in range(self.n_hash):
            nbytes, width = int(b[o]), int(b[o + 1])
            self.hash.append((nbytes, width, b[o + 2:o + 2 + width].astype(np.int64)))
            o += 2 + width
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

    def run(self, seed, nlanes, ncycles, stim0=0, state=None):
        """Lanes start from zero state (or ``state``: per-var values) and run
        ``ncycles`` cycles; returns hashes[nlanes]."""
        out = np.zeros(nlanes, np.uint64)
        for g0 in range(0, nlanes, 64):
            n = min(64, nlanes - g0)
            act = np.uint64((1 << n) - 1) if n < 64 else M64
            sids = np.arange(g0 + stim0, g0 + stim0 + 64, dtype=np.uint64)
            L = np.zeros(self.nslots, np.uint64)
            h = np.full(64, 0xCBF29CE484222325, np.uint64)
            for c in range(ncycles):
                for k, nb, s0 in self.inputs:
                    r = stim(seed, sids, c, int(k))
                    L[s0:s0 + nb] = to_planes(r, int(nb))
                self.run_phase(L, 0)
                st, dn = self.commit[:, 0], self.commit[:, 1]
                L[st] = (L[dn] & act) | (L[st] & ~act)
                self.run_phase(L, 1)
                for nbytes, width, refs in self.hash:
                    planes = np.array([self.ref(L, int(r)) for r in refs], np.uint64)
                    if width <= 64:
                        h = fnv_bytes(h, from_planes(planes), nbytes)
                    else:
                        for w0 in range(0, nbytes * 8, 32):
                            h = fnv_bytes(h, from_planes(planes[w0:w0 + 32]), 4)
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
    for c in range(ncycles):
        for k, nb, s0 in prog.inputs:
            L[s0:s0 + nb] = to_planes(stim(seed, sids, c, int(k)), int(nb))
        prog.run_phase(L, 0)
        st, dn = prog.commit[:, 0], prog.commit[:, 1]
        L[st] = (L[dn] & act) | (L[st] & ~act)
        prog.run_phase(L, 1)
    t = lane - g0
    out = {}
    for ph, wb in prog.fchunks:
        prog.run_phase(L, ph)
        for cls, loc, words, width, refs in wb:
            bits = [int((prog.ref(L, int(r)) >> np.uint64(t)) & np.uint64(1)) for r in refs]
            out[(cls, loc)] = sum(b << i for i, b in enumerate(bits))
    return out
