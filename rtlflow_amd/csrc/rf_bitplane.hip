// rf_bitplane.hip -- bit-plane evaluation of a design lowered to single-bit
// logic: the clock-cycle protocol (inputs, clk = 0 settle, clock edge,
// clk = 1 settle, output hash) for 64 stimuli per 64-bit word.
//
// Workgroup = RF_BP_THREADS threads for one group of 64 lanes. Wavefront 0
// does the per-lane work (stimulus generation, ballots into planes, output
// hashing, state load and write-back); every thread evaluates gates, one
// level of the program per barrier, with all values in LDS:
//
//   dst = s ? (x ^ ix) : (y ^ iy)        (a bit-select per 32-bit half)
//
// which covers AND, OR, XOR, NAND, NOR, XNOR, NOT and 2:1 MUX with the
// inversions folded into ix / iy.
#include "rf_bitplane.h"
#include "rf_stim.h"

namespace {

__device__ __forceinline__ uint64_t bp_ref(const uint64_t* L, uint32_t r) {
    const uint64_t v = L[r & 0x7fffffffu];
    return (r >> 31) ? ~v : v;
}

// 64 x 64 bit-matrix transpose across wavefront 0: thread j holds row j on
// entry and column j on exit (bit t of the result = bit j of thread t's
// row). Six block-swap stages (32, 16, ..., 1), each one cross-lane exchange.
__device__ __forceinline__ uint64_t bp_shfl_xor64(uint64_t v, uint32_t m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, (int)m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), (int)m, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t bp_transpose64(uint64_t x, uint32_t t) {
    const uint64_t lo_masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                  0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t s = 32u >> k;
        const uint64_t p = bp_shfl_xor64(x, s);
        const uint64_t lo = lo_masks[k];
        x = (t & s) ? ((x & ~lo) | ((p & ~lo) >> s)) : ((x & lo) | ((p & lo) << s));
    }
    return x;
}

// Bits [0, nbits) of every lane's v (wavefront 0, all 64 threads) become
// planes L[s0 .. s0 + nbits).
__device__ __forceinline__ void bp_to_planes(uint64_t* L, uint32_t s0, uint64_t v, uint32_t nbits,
                                             uint32_t t) {
    const uint64_t col = bp_transpose64(v, t);
    if (t < nbits) L[s0 + t] = col;
}

// This lane's bits of planes refs[0 .. nbits) (nbits <= 64).
__device__ __forceinline__ uint64_t bp_gather(const uint64_t* L, const uint32_t* refs, uint32_t nbits,
                                              uint32_t t) {
    const uint64_t row = t < nbits ? bp_ref(L, refs[t]) : 0ull;
    return bp_transpose64(row, t);
}

__device__ __forceinline__ void bp_phase(uint64_t* L, const uint32_t* lev, const uint2* gates,
                                         uint32_t nlev, uint32_t tid) {
    for (uint32_t l = 0; l < nlev; ++l) {
        const uint32_t g1 = lev[l + 1];
        for (uint32_t g = lev[l] + tid; g < g1; g += RF_BP_THREADS) {
            const uint2 w = gates[g];
            const uint32_t dst = w.x & 0x7fffu;
            const uint32_t s = (w.x >> 15) & 0x7fffu;
            const uint32_t xs = ((w.x >> 30) | (w.y << 2)) & 0x7fffu;
            const uint32_t ys = (w.y >> 14) & 0x7fffu;
            const uint64_t ix = 0ull - (uint64_t)((w.y >> 13) & 1u);
            const uint64_t iy = 0ull - (uint64_t)((w.y >> 29) & 1u);
            const uint64_t S = L[s];
            const uint64_t X = L[xs] ^ ix;
            const uint64_t Y = L[ys] ^ iy;
            L[dst] = (S & X) | (~S & Y);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t bp_load(const rf_args& a, uint32_t cls, uint64_t slot, uint32_t lane) {
    const size_t i = (size_t)slot * a.npad + lane;
    switch (cls) {
        case RF_CLS_C: return a.c[i];
        case RF_CLS_S: return a.s[i];
        case RF_CLS_I: return a.i[i];
        default: return a.q[i];
    }
}

__device__ __forceinline__ void bp_store(const rf_args& a, uint32_t cls, uint64_t slot, uint32_t lane,
                                         uint64_t v) {
    const size_t i = (size_t)slot * a.npad + lane;
    switch (cls) {
        case RF_CLS_C: a.c[i] = (uint8_t)v; break;
        case RF_CLS_S: a.s[i] = (uint16_t)v; break;
        case RF_CLS_I: a.i[i] = (uint32_t)v; break;
        default: a.q[i] = v; break;
    }
}

__global__ void __launch_bounds__(RF_BP_THREADS) rf_bp_kernel(rf_bp_params P) {
    extern __shared__ uint64_t L[];
    __shared__ uint64_t s_act;
    const rf_args& a = P.a;
    const uint32_t tid = threadIdx.x;
    const uint32_t t = tid & 63u;
    const bool w0 = tid < 64u;
    const uint32_t lane = blockIdx.x * 64u + t;
    const bool inb = lane < a.nlanes;
    const uint32_t* prog = P.prog;
    bool act = false;
    uint64_t h = 0, sid = 0;
    uint32_t cyc = 0;
    if (w0) {
        if (inb) {
            act = a.done[lane] == 0;
            h = a.hash[lane];
            cyc = a.cyc[lane];
            sid = a.stim[lane];
        }
        const uint64_t am = __ballot((int)act);
        if (t == 0) {
            s_act = am;
            L[0] = 0;
        }
        const uint32_t* sl = prog + P.off_state;
        for (uint32_t k = 0; k < P.n_state; ++k, sl += 6) {
            const uint32_t cls = sl[0], words = sl[3], width = sl[4], s0 = sl[5];
            const uint64_t loc = (uint64_t)sl[1] | ((uint64_t)sl[2] << 32);
            if (width <= 64) {
                const uint64_t v = inb ? bp_load(a, cls, loc, lane) : 0;
                bp_to_planes(L, s0, v, width, t);
            } else {
                for (uint32_t w = 0; w < words; ++w) {
                    const uint64_t v = inb ? a.i[(size_t)(loc + w) * a.npad + lane] : 0;
                    const uint32_t nb = min(32u, width - 32u * w);
                    bp_to_planes(L, s0 + 32u * w, v, nb, t);
                }
            }
        }
    }
    const uint32_t* lev[3];
    const uint2* gates[3];
    for (int k = 0; k < 3; ++k) {
        lev[k] = prog + P.off_lev[k];
        gates[k] = reinterpret_cast<const uint2*>(prog + P.off_gates[k]);
    }
    const uint32_t* commit = prog + P.off_commit;
    for (uint32_t n = 0; n < a.ncycles; ++n) {
        if (w0) {
            const uint32_t* ic = prog + P.off_in;
            for (uint32_t k = 0; k < P.n_in; ++k, ic += 3) {
                const uint64_t r = rf_stim(a.seed, sid, (uint64_t)(cyc + n), ic[0]);
                bp_to_planes(L, ic[2], r, ic[1], t);
            }
        }
        __syncthreads();
        bp_phase(L, lev[0], gates[0], P.nlev[0], tid);
        {
            const uint64_t am = s_act;
            for (uint32_t i = tid; i < P.n_commit; i += RF_BP_THREADS) {
                const uint32_t st = commit[2 * i], dn = commit[2 * i + 1];
                L[st] = (L[dn] & am) | (L[st] & ~am);
            }
        }
        __syncthreads();
        bp_phase(L, lev[1], gates[1], P.nlev[1], tid);
        if (w0) {
            const uint32_t* hp = prog + P.off_hash;
            uint64_t hn = h;
            for (uint32_t k = 0; k < P.n_hash; ++k) {
                const uint32_t nbytes = hp[0], width = hp[1];
                const uint32_t* refs = hp + 2;
                if (width <= 64) {
                    hn = rf_fnv_bytes(hn, bp_gather(L, refs, width, t), (int)nbytes);
                } else {
                    for (uint32_t w = 0; 4u * w < nbytes; ++w) {
                        const uint32_t nb = min(32u, width - 32u * w);
                        hn = rf_fnv_bytes(hn, bp_gather(L, refs + 32u * w, nb, t), 4);
                    }
                }
                hp += 2 + width;
            }
            if (act) h = hn;
        }
    }
    // settled state after the last cycle, written back to the lane-minor arrays
    __syncthreads();
    bp_phase(L, lev[2], gates[2], P.nlev[2], tid);
    if (w0 && act) {
        const uint32_t* wp = prog + P.off_wb;
        for (uint32_t k = 0; k < P.n_wb; ++k) {
            const uint32_t cls = wp[0], words = wp[3], width = wp[4];
            const uint64_t loc = (uint64_t)wp[1] | ((uint64_t)wp[2] << 32);
            const uint32_t* refs = wp + 5;
            if (width <= 64) {
                bp_store(a, cls, loc, lane, bp_gather(L, refs, width, t));
            } else {
                for (uint32_t w = 0; w < words; ++w) {
                    const uint32_t nb = min(32u, width - 32u * w);
                    a.i[(size_t)(loc + w) * a.npad + lane] = (uint32_t)bp_gather(L, refs + 32u * w, nb, t);
                }
            }
            wp += 5 + width;
        }
        a.hash[lane] = h;
        a.cyc[lane] = cyc + a.ncycles;
    }
}

}  // namespace

int rf_bp_parse(const uint32_t* b, uint64_t words, rf_bp_params* p, const char** why) {
    if (words < RF_BP_HDR_WORDS || b[0] != RF_BP_MAGIC || b[1] != RF_BP_VERSION) {
        *why = "bad bit-plane program header";
        return -1;
    }
    p->nslots = b[2];
    p->n_in = b[3];
    p->n_state = b[4];
    p->n_commit = b[5];
    p->n_hash = b[6];
    p->n_wb = b[7];
    for (int k = 0; k < 3; ++k) p->nlev[k] = b[8 + k];
    const uint32_t ngates[3] = {b[11], b[12], b[13]};
    if (b[14] != words) {
        *why = "bit-plane program size mismatch";
        return -1;
    }
    auto ev = [](uint64_t x) { return x + (x & 1); };
    uint64_t o = ev(RF_BP_HDR_WORDS);
    p->off_in = (uint32_t)o;
    o += 3ull * p->n_in;
    o = ev(o);
    p->off_state = (uint32_t)o;
    o += 6ull * p->n_state;
    o = ev(o);
    p->off_commit = (uint32_t)o;
    o += 2ull * p->n_commit;
    for (int k = 0; k < 3; ++k) {
        o = ev(o);
        p->off_lev[k] = (uint32_t)o;
        o += p->nlev[k] + 1ull;
        o = ev(o);
        p->off_gates[k] = (uint32_t)o;
        o += 2ull * ngates[k];
    }
    o = ev(o);
    p->off_hash = (uint32_t)o;
    for (uint32_t k = 0; k < p->n_hash && o + 2 <= words; ++k) o += 2ull + b[o + 1];
    o = ev(o);
    p->off_wb = (uint32_t)o;
    for (uint32_t k = 0; k < p->n_wb && o + 5 <= words; ++k) o += 5ull + b[o + 4];
    if (ev(o) != words) {
        *why = "bit-plane program sections do not add up";
        return -1;
    }
    if (p->nslots == 0 || p->nslots > 0x7fffu || (uint64_t)p->nslots * 8u > 160u * 1024u - 64u) {
        *why = "bit-plane program needs more LDS than a workgroup has";
        return -1;
    }
    // every slot index in the program must be below nslots
    for (int k = 0; k < 3; ++k) {
        const uint32_t* lv = b + p->off_lev[k];
        if (lv[0] != 0 || lv[p->nlev[k]] != ngates[k]) {
            *why = "bit-plane level table inconsistent";
            return -1;
        }
        for (uint32_t l = 0; l < p->nlev[k]; ++l)
            if (lv[l + 1] < lv[l]) {
                *why = "bit-plane level table not monotonic";
                return -1;
            }
        const uint32_t* g = b + p->off_gates[k];
        for (uint32_t i = 0; i < ngates[k]; ++i) {
            const uint64_t w = (uint64_t)g[2 * i] | ((uint64_t)g[2 * i + 1] << 32);
            const uint32_t f[4] = {(uint32_t)(w & 0x7fff), (uint32_t)((w >> 15) & 0x7fff),
                                   (uint32_t)((w >> 30) & 0x7fff), (uint32_t)((w >> 46) & 0x7fff)};
            for (uint32_t x : f)
                if (x >= p->nslots) {
                    *why = "bit-plane gate slot out of range";
                    return -1;
                }
        }
    }
    for (uint32_t k = 0; k < p->n_in; ++k) {
        const uint32_t* ic = b + p->off_in + 3 * k;
        if (ic[1] > 64 || ic[2] + ic[1] > p->nslots) {
            *why = "bit-plane input slot out of range";
            return -1;
        }
    }
    for (uint32_t k = 0; k < p->n_state; ++k) {
        const uint32_t* sl = b + p->off_state + 6 * k;
        if (sl[0] > 3 || sl[4] == 0 || sl[5] + sl[4] > p->nslots ||
            (sl[4] > 64 && (sl[0] != RF_CLS_I || sl[3] * 32u < sl[4] || sl[3] * 32u >= sl[4] + 32u))) {
            *why = "bit-plane state entry out of range";
            return -1;
        }
    }
    for (uint32_t k = 0; k < p->n_commit; ++k) {
        const uint32_t* c = b + p->off_commit + 2 * k;
        if (c[0] >= p->nslots || c[1] >= p->nslots) {
            *why = "bit-plane commit slot out of range";
            return -1;
        }
    }
    uint64_t q = p->off_hash;
    for (uint32_t k = 0; k < p->n_hash; ++k) {
        const uint32_t nbytes = b[q], width = b[q + 1];
        if (width == 0 || (width <= 64 && (nbytes > 8 || nbytes * 8u < width)) ||
            (width > 64 && (nbytes % 4u || nbytes * 8u < width || nbytes * 8u >= width + 32u))) {
            *why = "bit-plane hash entry malformed";
            return -1;
        }
        for (uint32_t i = 0; i < width; ++i)
            if ((b[q + 2 + i] & 0x7fffffffu) >= p->nslots) {
                *why = "bit-plane hash slot out of range";
                return -1;
            }
        q += 2 + width;
    }
    q = p->off_wb;
    for (uint32_t k = 0; k < p->n_wb; ++k) {
        const uint32_t cls = b[q], words_ = b[q + 3], width = b[q + 4];
        if (cls > 3 || width == 0 || (width > 64 && (cls != RF_CLS_I || words_ * 32u < width ||
                                                      words_ * 32u >= width + 32u))) {
            *why = "bit-plane write-back entry malformed";
            return -1;
        }
        for (uint32_t i = 0; i < width; ++i)
            if ((b[q + 5 + i] & 0x7fffffffu) >= p->nslots) {
                *why = "bit-plane write-back slot out of range";
                return -1;
            }
        q += 5 + width;
    }
    return 0;
}

hipError_t rf_bp_launch(const rf_bp_params& p, hipStream_t stream) {
    const size_t shmem = (size_t)p.nslots * 8u;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&rf_bp_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const uint32_t grid = (p.a.nlanes + 63u) / 64u;
    hipLaunchKernelGGL(rf_bp_kernel, dim3(grid), dim3(RF_BP_THREADS), shmem, stream, p);
    return hipGetLastError();
}
