"""NumPy model of a gate-net program blob (include/rf_gatenet.h), for CPU
tests: the full eval by levels, the rising edge (flip-flops, RAM ports), the
output cone's eval, the per-cycle hash -- or, with ``order="mtasks"``, both
evals through the mtask-ordered records. Test infrastructure: checks the
program the host side writes against the oracle without a GPU."""
from __future__ import annotations

import numpy as np

H = {k: i for i, k in enumerate(
    "MAGIC VERSION NODES PI FF RAMBITS PO RAMS DEPTH WIDTH ABITS LEV_FULL LEV_CONE REC_FULL REC_CONE "
    "MTASKS OFF_LEV_FULL OFF_LEV_CONE OFF_FF OFF_PO OFF_RAM OFF_REC_FULL OFF_REC_CONE OFF_MT OFF_PREDS "
    "OFF_REC_MT REC_MT TOTAL".split())}
FNV_BASIS = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
M64 = (1 << 64) - 1


def _splitmix(x):
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _stim(seed, s, c, k):
    return _splitmix(seed ^ ((s * 0x9E3779B97F4A7C15) & M64) ^ ((c << 8) & M64) ^ k)


def _eval(x, rec):
    for r in rec:
        ty = r[0] >> 29
        dst = r[0] & ((1 << 29) - 1)
        a, b, c = x[:, r[1]], x[:, r[2]], x[:, r[3]]
        if ty == 0:
            v = a & b
        elif ty == 1:
            v = a | b
        elif ty == 2:
            v = a ^ b
        elif ty == 3:
            v = (a & b) ^ 1
        else:
            v = np.where(a != 0, b, c)
        x[:, dst] = v


def run(prog: np.ndarray, seed: int, budgets, order: str = "levels") -> list:
    h = prog
    g = lambda k: int(h[H[k]])
    n = len(budgets)
    nodes, npi, nff, npo = g("NODES"), g("PI"), g("FF"), g("PO")
    nr, depth, width, ab = g("RAMS"), g("DEPTH"), g("WIDTH"), g("ABITS")
    rec_full = h[g("OFF_REC_FULL"):g("OFF_REC_FULL") + 4 * g("REC_FULL")].reshape(-1, 4).astype(np.int64)
    rec_cone = h[g("OFF_REC_CONE"):g("OFF_REC_CONE") + 4 * g("REC_CONE")].reshape(-1, 4).astype(np.int64)
    rec_mt = h[g("OFF_REC_MT"):g("OFF_REC_MT") + 4 * g("REC_MT")].reshape(-1, 4).astype(np.int64)
    ffd = h[g("OFF_FF"):g("OFF_FF") + nff].astype(np.int64)
    po = h[g("OFF_PO"):g("OFF_PO") + npo].astype(np.int64)
    ram = h[g("OFF_RAM"):g("OFF_RAM") + nr * (ab + 1 + width)].reshape(nr, ab + 1 + width).astype(np.int64) \
        if nr else np.zeros((0, 0), np.int64)
    x = np.zeros((n, nodes), np.uint8)
    mem = np.zeros((n, nr, depth), np.uint64)
    hashes = [FNV_BASIS] * n
    nb = 1 if npo <= 8 else 2 if npo <= 16 else 4 if npo <= 32 else 8
    first_eval = rec_mt if order == "mtasks" else rec_full
    second_eval = rec_mt if order == "mtasks" else rec_cone
    for c in range(int(max(budgets))):
        for lane in range(n):
            for j in range((npi + 63) // 64):
                w = _stim(seed, lane, c, j)
                for i in range(64 * j, min(npi, 64 * j + 64)):
                    x[lane, i] = (w >> (i - 64 * j)) & 1
        _eval(x, first_eval)
        d = x[:, ffd].copy()
        for r in range(nr):
            addr = np.zeros(n, np.int64)
            for k in range(ab):
                addr |= x[:, ram[r, k]].astype(np.int64) << k
            we = x[:, ram[r, ab]]
            wd = np.zeros(n, np.uint64)
            for k in range(width):
                wd |= x[:, ram[r, ab + 1 + k]].astype(np.uint64) << np.uint64(k)
            for lane in range(n):
                a = int(addr[lane])
                old = int(mem[lane, r, a]) if a < depth else 0
                if we[lane] and a < depth:
                    mem[lane, r, a] = wd[lane]
                for k in range(width):
                    x[lane, npi + nff + r * width + k] = (old >> k) & 1
        x[:, npi:npi + nff] = d
        _eval(x, second_eval)
        for lane in range(n):
            if c < budgets[lane]:
                v = 0
                for k in range(npo):
                    v |= int(x[lane, po[k]]) << k
                hv = hashes[lane]
                for i in range(nb):
                    hv = ((hv ^ ((v >> (8 * i)) & 0xFF)) * FNV_PRIME) & M64
                hashes[lane] = hv
    return hashes
