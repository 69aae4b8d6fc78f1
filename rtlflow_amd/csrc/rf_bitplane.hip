// rf_bitplane.hip -- bit-plane evaluation of a design lowered to single-bit
// logic: the clock-cycle protocol (inputs, clk = 0 settle, clock edge,
// clk = 1 settle, output hash) for 64 stimuli per 64-bit word.
//
// One workgroup (128 to 1,024 threads) per group of 64 lanes. Wavefronts 0
// and 1 do the per-cycle lane work (stimulus generation, transposes into
// planes, output hashing); synchronous memory ports go to the other
// wavefronts; all wavefronts share the state load and write-back at launch
// start and end. Every thread evaluates gates, one level of the program per
// barrier, with all values in LDS:
//
//   dst = s ? (x ^ ix) : (y ^ iy)        (a bit-select per 32-bit half)
//
// which covers AND, OR, XOR, NAND, NOR, XNOR, NOT and 2:1 MUX with the
// inversions folded into ix / iy.
#include <algorithm>

#include "rf_bitplane.h"
#include "rf_stim.h"
#include "rf_transpose.h"

namespace {

typedef __attribute__((address_space(3))) uint64_t lds_u64;

__device__ __forceinline__ uint64_t bp_ref(const uint64_t* L, uint32_t r) {
    const uint64_t v = L[r & 0x7fffffffu];
    return (r >> 31) ? ~v : v;
}

// Bits [0, nbits) of every lane's v (wavefront 0, all 64 threads) become
// planes L[s0 .. s0 + nbits).
__device__ __forceinline__ void bp_to_planes(uint64_t* L, uint32_t s0, uint64_t v, uint32_t nbits,
                                             uint32_t t) {
    const uint64_t col = bp_transpose64(v, t);
    if (t < nbits) L[s0 + t] = col;
}

// This lane's bits of planes refs[0 .. nbits) (nbits <= 64).
template <typename TP>
__device__ __forceinline__ uint64_t bp_gather(const uint64_t* L, TP refs, uint32_t nbits, uint32_t t) {
    const uint64_t row = t < nbits ? bp_ref(L, refs[t]) : 0ull;
    return bp_transpose64(row, t);
}
// Low bits of 16 bytes (each 0 or 1) as 16 bits, byte i to bit i: per
// 32-bit word, the multiply moves bit 8i to bit 24 + i without carries.
__device__ __forceinline__ uint32_t bp_pack16(uint4 v) {
    const auto p4 = [](uint32_t w) { return ((w & 0x01010101u) * 0x01020408u) >> 24; };
    return p4(v.x) | (p4(v.y) << 4) | (p4(v.z) << 8) | (p4(v.w) << 12);
}
// 16 bits as 16 bytes of 0 or 1, bit i to byte i: per 4 bits, the multiply
// moves bit i to bit 8i without carries.
__device__ __forceinline__ uint4 bp_unpack16(uint32_t b) {
    const auto u4 = [](uint32_t x) { return ((x & 15u) * 0x00204081u) & 0x01010101u; };
    return make_uint4(u4(b), u4(b >> 4), u4(b >> 8), u4(b >> 12));
}
// Single-bit variables (class C, one byte per lane) of a 64-lane group,
// four threads per variable, each moving its 16 lanes' bytes with one
// 16-byte access: from the lane-minor array into planes (load), or from
// planes back (store). ``tab`` holds 6-word entries (loc in words 1-2, the
// plane's slot or reference in word 5); lanes outside ``lanes`` (the load:
// lanes that exist, the store: lanes that write back) are zero in the
// planes, resp. left untouched in memory. U entries per thread are in
// flight at once.
constexpr uint32_t kBpU = 4;
template <int NT>
__device__ __forceinline__ void bp_load_bits(uint64_t* L, const rf_args& a, const uint32_t* tab, uint32_t n,
                                             uint32_t grp, uint64_t lanes, uint32_t tid) {
    constexpr uint32_t Q = NT / 4;
    const uint32_t q = tid & 3u;
    const size_t col = (size_t)grp * 64u + 16u * q;
    const uint32_t qm = (uint32_t)(lanes >> (16u * q)) & 0xffffu;
    for (uint32_t k0 = tid >> 2; k0 < n; k0 += kBpU * Q) {
        uint64_t loc[kBpU];
        uint32_t sl[kBpU];
#pragma unroll
        for (uint32_t j = 0; j < kBpU; ++j) {
            const uint32_t* e = tab + 6u * min(k0 + j * Q, n - 1u);
            loc[j] = (uint64_t)e[1] | ((uint64_t)e[2] << 32);
            sl[j] = e[5];
        }
        uint4 v[kBpU];
#pragma unroll
        for (uint32_t j = 0; j < kBpU; ++j) v[j] = *reinterpret_cast<const uint4*>(a.c + loc[j] * a.npad + col);
#pragma unroll
        for (uint32_t j = 0; j < kBpU; ++j)
            if (k0 + j * Q < n) reinterpret_cast<uint16_t*>(L + sl[j])[q] = (uint16_t)(bp_pack16(v[j]) & qm);
    }
}
template <int NT>
__device__ __forceinline__ void bp_store_bits(const uint64_t* L, const rf_args& a, const uint32_t* tab, uint32_t n,
                                              uint32_t grp, uint64_t lanes, uint32_t tid) {
    constexpr uint32_t Q = NT / 4;
    const uint32_t q = tid & 3u;
    const size_t col = (size_t)grp * 64u + 16u * q;
    const uint32_t qm = (uint32_t)(lanes >> (16u * q)) & 0xffffu;
    if (!qm) return;
    for (uint32_t k0 = tid >> 2; k0 < n; k0 += kBpU * Q) {
        uint64_t loc[kBpU];
        uint32_t ref[kBpU];
#pragma unroll
        for (uint32_t j = 0; j < kBpU; ++j) {
            const uint32_t* e = tab + 6u * min(k0 + j * Q, n - 1u);
            loc[j] = (uint64_t)e[1] | ((uint64_t)e[2] << 32);
            ref[j] = e[5];
        }
#pragma unroll
        for (uint32_t j = 0; j < kBpU; ++j) {
            if (k0 + j * Q >= n) continue;
            const uint32_t b = (uint32_t)(bp_ref(L, ref[j]) >> (16u * q)) & 0xffffu;
            uint8_t* dst = a.c + loc[j] * a.npad + col;
            if (qm == 0xffffu) {
                *reinterpret_cast<uint4*>(dst) = bp_unpack16(b);
            } else {
                for (uint32_t i = 0; i < 16u; ++i)
                    if ((qm >> i) & 1u) dst[i] = (uint8_t)((b >> i) & 1u);
            }
        }
    }
}

// Gate record, slot form: x = dst | s << 16; y = xs | ix << 15 | ys << 16 |
// iy << 31. Byte form (BA, programs of at most 8,192 slots): every field is
// a slot's LDS byte offset (slot * 8), with the inversion flags in bit 0 of
// the x and y fields: x = dst8 | s8 << 16; y = (xs8 | ix) | (ys8 | iy) << 16.
// The byte form needs no shift per operand address.
// The gate on decoded byte-form fields (LDS byte offsets; ix, iy 0 or ~0).
__device__ __forceinline__ void bp_gate_bytes(uint32_t dst, uint32_t s, uint32_t xs, uint32_t ys, uint32_t ix,
                                              uint32_t iy) {
    const uint64_t S = *reinterpret_cast<const lds_u64*>((size_t)s);
    const uint64_t X = *reinterpret_cast<const lds_u64*>((size_t)xs);
    const uint64_t Y = *reinterpret_cast<const lds_u64*>((size_t)ys);
    const uint32_t xl = (uint32_t)X ^ ix, xh = (uint32_t)(X >> 32) ^ ix;
    const uint32_t yl = (uint32_t)Y ^ iy, yh = (uint32_t)(Y >> 32) ^ iy;
    const uint32_t sl = (uint32_t)S, sh = (uint32_t)(S >> 32);
    *reinterpret_cast<lds_u64*>((size_t)dst) =
        ((uint64_t)((sh & xh) | (~sh & yh)) << 32) | (uint64_t)((sl & xl) | (~sl & yl));
}

// A register-resident byte-form record, decoded by an opaque instruction
// sequence: the compiler can neither hoist the six decoded fields out of
// the cycle loop (six registers per record instead of two) nor needs a copy
// of the record to keep it intact. The empty record (x = 0xffffffff, no
// valid destination offset 0xffff) is tested on the decoded field for the
// same reason: a test of the record itself would be hoisted as a lane mask
// per record.
__device__ __forceinline__ void bp_gate_bytes_live(uint2 w) {
    uint32_t dst, s, xs, ys, ix, iy;
    asm volatile(
        "v_and_b32 %0, 0xffff, %6\n\t"
        "v_lshrrev_b32 %1, 16, %6\n\t"
        "v_and_b32 %2, 0xfff8, %7\n\t"
        "v_and_b32_sdwa %3, %7, %8 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:DWORD\n\t"
        "v_bfe_i32 %4, %7, 0, 1\n\t"
        "v_bfe_i32 %5, %7, 16, 1"
        : "=&v"(dst), "=&v"(s), "=&v"(xs), "=&v"(ys), "=&v"(ix), "=&v"(iy)
        : "v"(w.x), "v"(w.y), "s"(0xfff8u));
    if (dst != 0xffffu) bp_gate_bytes(dst, s, xs, ys, ix, iy);
}

// The slot form's register-resident record, decoded the same way into LDS
// byte offsets (the planes start at LDS address 0): dst and s are the two
// halves of x, shifted; xs and ys drop the inversion flags (bits 15 and 31
// of y). The empty record's destination half 0xffff becomes 0x7fff8, which
// no slot has (at most 0x7fff slots).
__device__ __forceinline__ void bp_gate_slots_live(uint2 w) {
    uint32_t dst, s, xs, ys, ix, iy;
    asm volatile(
        "v_lshlrev_b32_sdwa %0, %8, %6 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0\n\t"
        "v_lshlrev_b32_sdwa %1, %8, %6 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1\n\t"
        "v_lshlrev_b32 %2, 3, %7\n\t"
        "v_and_b32 %2, %9, %2\n\t"
        "v_lshrrev_b32 %3, 13, %7\n\t"
        "v_and_b32 %3, %9, %3\n\t"
        "v_bfe_i32 %4, %7, 15, 1\n\t"
        "v_ashrrev_i32 %5, 31, %7"
        : "=&v"(dst), "=&v"(s), "=&v"(xs), "=&v"(ys), "=&v"(ix), "=&v"(iy)
        : "v"(w.x), "v"(w.y), "s"(3u), "s"(0x3fff8u));
    if (dst != 0x7fff8u) bp_gate_bytes(dst, s, xs, ys, ix, iy);
}

template <bool BA>
__device__ __forceinline__ void bp_gate(uint64_t* L, uint2 w) {
    if constexpr (BA) {
        // the planes are the kernel's only LDS allocation (no static
        // __shared__ variable), so they start at LDS address 0 and a
        // record's byte offsets are LDS addresses
        (void)L;
        const uint32_t dst = w.x & 0xffffu;
        const uint32_t s = w.x >> 16;
        const uint32_t xs = w.y & 0xfff8u;
        const uint32_t ys = (w.y >> 16) & 0xfff8u;
        const uint32_t ix = (uint32_t)((int32_t)(w.y << 31) >> 31);   // 0 or ~0
        const uint32_t iy = (uint32_t)((int32_t)(w.y << 15) >> 31);
        bp_gate_bytes(dst, s, xs, ys, ix, iy);
        return;
    }
    const uint32_t dst = w.x & 0xffffu;
    const uint32_t s = w.x >> 16;
    const uint32_t xs = w.y & 0x7fffu;
    const uint32_t ys = (w.y >> 16) & 0x7fffu;
    const uint32_t ix = (uint32_t)((int32_t)(w.y << 16) >> 31);   // 0 or ~0
    const uint32_t iy = (uint32_t)((int32_t)w.y >> 31);
    const uint64_t S = L[s];
    const uint64_t X = L[xs];
    const uint64_t Y = L[ys];
    const uint32_t xl = (uint32_t)X ^ ix, xh = (uint32_t)(X >> 32) ^ ix;
    const uint32_t yl = (uint32_t)Y ^ iy, yh = (uint32_t)(Y >> 32) ^ iy;
    const uint32_t sl = (uint32_t)S, sh = (uint32_t)(S >> 32);
    L[dst] = ((uint64_t)((sh & xh) | (~sh & yh)) << 32) | (uint64_t)((sl & xl) | (~sl & yl));
}

// One phase: level l's gates are spread over the workgroup, a barrier after
// each level. Each thread's first descriptor of the next level is loaded
// before the current level is evaluated, so its global-memory latency
// overlaps the current level and the barrier.
template <int NT, bool BA>
__device__ __forceinline__ void bp_phase(uint64_t* L, const uint32_t* lev, const uint2* gates,
                                         uint32_t nlev, uint32_t tid) {
    if (nlev == 0) return;
    uint32_t lo = lev[0], hi = lev[1];
    uint2 nxt = (lo + tid < hi) ? gates[lo + tid] : make_uint2(0u, 0u);
    for (uint32_t l = 0; l < nlev; ++l) {
        const uint2 cur = nxt;
        const uint32_t g0 = lo + tid;
        const uint32_t g1 = hi;
        if (l + 1 < nlev) {
            lo = hi;
            hi = lev[l + 2];
            if (lo + tid < hi) nxt = gates[lo + tid];
        }
        if (g0 < g1) {
            bp_gate<BA>(L, cur);
            for (uint32_t g = g0 + NT; g < g1; g += NT) bp_gate<BA>(L, gates[g]);
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

// The per-cycle tables (input chunks, hashed outputs of program M, commit
// pairs of program M) are read from global memory, or from the copy the
// kernel makes in LDS after the planes (lds_u32 pointers) when the launch
// reserves room for it.
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// Inputs of one cycle into their planes (wavefront 0): set A, read by the
// clk = 0 settle. Chunk k of the table is input word k (the generator's or
// the input stream's).
// ``live``: the lane runs that cycle (inactive lanes and lanes past their
// budget still fill their plane bits, which nothing reads). IO: the launch
// may read an input stream (otherwise the generator only, with no stream
// code in the cycle loop).
template <bool IO, typename TP>
__device__ __forceinline__ void bp_inputs(uint64_t* L, TP ic, uint32_t n_in, const rf_args& a, uint64_t sid,
                                          uint32_t cyc, uint32_t t, bool live) {
    for (uint32_t k = 0; k < n_in; ++k, ic += 4) {
        uint64_t r;
        if constexpr (IO)
            r = rf_input(a, sid, (uint64_t)cyc, ic[0], k, live);
        else
            r = rf_stim(a.seed, sid, (uint64_t)cyc, ic[0]);
        bp_to_planes(L, ic[2], r, ic[1], t);
    }
}

// Input set A -> set B (the clk = 1 settle of the same cycle reads set B
// while set A already holds the next cycle's inputs); wavefront 0.
template <typename TP>
__device__ __forceinline__ void bp_copy_inputs(uint64_t* L, TP ic, uint32_t n_in, uint32_t t) {
    for (uint32_t k = 0; k < n_in; ++k, ic += 4)
        if (t < ic[1]) L[ic[3] + t] = L[ic[2] + t];
}

// Fold this cycle's outputs into the lane's hash (wavefront 1); ``hp``: the
// output-slot table of program M or B. With ``cap`` the outputs' 64-bit
// words also go to the output capture as stimulus ``sid``'s cycle ``c``
// (IO launches only).
template <bool IO, typename TP>
__device__ __forceinline__ uint64_t bp_hash(const uint64_t* L, TP hp, uint32_t n_hash, uint64_t h, uint32_t t,
                                            const rf_args& a, bool cap, uint64_t sid, uint64_t c) {
    if constexpr (!IO) cap = false;
    uint32_t n = 0;
    for (uint32_t k = 0; k < n_hash; ++k) {
        const uint32_t nbytes = hp[0], width = hp[1];
        const TP refs = hp + 2;
        if (width <= 64) {
            const uint64_t v = bp_gather(L, refs, width, t);
            h = rf_fnv_bytes(h, v, (int)nbytes);
            if (cap) rf_capture(a, sid, c, n, v);
            ++n;
        } else {
            uint64_t lo = 0;
            for (uint32_t w = 0; 4u * w < nbytes; ++w) {
                const uint32_t nb = min(32u, width - 32u * w);
                const uint64_t v = bp_gather(L, refs + 32u * w, nb, t);
                h = rf_fnv_bytes(h, v, 4);
                if (w & 1u) {
                    if (cap) rf_capture(a, sid, c, n, lo | (v << 32));
                    ++n;
                } else {
                    lo = v;
                }
            }
            if ((nbytes >> 2) & 1u) {
                if (cap) rf_capture(a, sid, c, n, lo);
                ++n;
            }
        }
        hp += 2 + width;
    }
    return h;
}

// Memory table entry: {design memory, cls, layout, loc_lo, loc_hi, row
// bytes, depth, width}; element e of lane ``lane`` lives where the per-lane
// kernels keep it (element-major: slot loc + e of its class array;
// lane-major: byte row ``lane`` of the region at loc * npad).
__device__ __forceinline__ uint64_t bp_mem_load(const rf_args& a, const uint32_t* me, uint32_t lane,
                                                uint64_t e) {
    const uint64_t loc = (uint64_t)me[3] | ((uint64_t)me[4] << 32);
    if (!me[2]) return bp_load(a, me[1], loc + e, lane);
    const uint8_t* row = a.mem + loc * a.npad + (size_t)lane * me[5];
    switch (me[1]) {
        case RF_CLS_C: return row[e];
        case RF_CLS_S: return reinterpret_cast<const uint16_t*>(row)[e];
        case RF_CLS_I: return reinterpret_cast<const uint32_t*>(row)[e];
        default: return reinterpret_cast<const uint64_t*>(row)[e];
    }
}

__device__ __forceinline__ void bp_mem_store(const rf_args& a, const uint32_t* me, uint32_t lane, uint64_t e,
                                             uint64_t v) {
    const uint64_t loc = (uint64_t)me[3] | ((uint64_t)me[4] << 32);
    if (!me[2]) {
        bp_store(a, me[1], loc + e, lane, v);
        return;
    }
    uint8_t* row = a.mem + loc * a.npad + (size_t)lane * me[5];
    switch (me[1]) {
        case RF_CLS_C: row[e] = (uint8_t)v; break;
        case RF_CLS_S: reinterpret_cast<uint16_t*>(row)[e] = (uint16_t)v; break;
        case RF_CLS_I: reinterpret_cast<uint32_t*>(row)[e] = (uint32_t)v; break;
        default: reinterpret_cast<uint64_t*>(row)[e] = v; break;
    }
}

// Memory ports at the clock edge, after program ``which`` (0: A0, 1: M).
// Memory mi is served by wavefront (2 + mi) mod NW, all 64 threads of it
// (lane t of the group is thread t): per port in list order (a memory's
// reads before its writes) each lane gathers its address (and guard and
// data) from the planes; a read loads the lane's old word and transposes
// the words into the target state planes; a write stores when the guard is
// set, the address is in range and the lane is active. Reads and writes of
// one memory stay in one thread per lane, in program order.
__device__ __forceinline__ void bp_mem_ports(uint64_t* L, const rf_bp_params& P, uint32_t which,
                                             uint32_t wave, uint32_t nw, uint32_t t, uint32_t lane,
                                             bool inb, bool act) {
    const uint32_t* pp = P.prog + P.off_ports[which];
    const uint32_t n = pp[1];
    const uint32_t* mt = P.prog + P.off_mem;
    pp += 2;
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t kind = pp[0], mi = pp[1], na = pp[2], n2 = pp[3];
        if ((2u + mi) % nw == wave) {
            const uint32_t* me = mt + RF_BP_MEM_WORDS * mi;
            const uint32_t depth = me[6];
            if (kind == 0) {
                const uint64_t e = bp_gather(L, pp + 4, na, t);
                const uint64_t v = (inb && e < depth) ? bp_mem_load(P.a, me, lane, e) : 0ull;
                const uint64_t col = bp_transpose64(v, t);
                const uint32_t* tg = pp + 4 + na;
                for (uint32_t j = 0; j < n2; ++j)
                    if ((tg[j] >> 24) == t) L[tg[j] & 0xffffffu] = col;
            } else {
                const bool g = (bp_ref(L, pp[4]) >> t) & 1ull;
                const uint64_t e = bp_gather(L, pp + 5, na, t);
                const uint64_t v = bp_gather(L, pp + 5 + na, n2, t);
                if (act && g && e < depth) bp_mem_store(P.a, me, lane, e, v);
            }
        }
        pp += (kind == 0 ? 4u : 5u) + na + n2;
    }
}

// Commit: state <- next state (the clock edge) for the state bits the
// program could not write in place. Unmasked: a lane that is not active
// (done, or past the last stimulus) never writes its state or hash back, so
// its planes may change freely during the launch.
template <int NT>
__device__ __forceinline__ void bp_commit(uint64_t* L, const uint32_t* c, uint32_t n, uint32_t tid) {
    for (uint32_t i = tid; i < n; i += NT) L[c[2 * i]] = L[c[2 * i + 1]];
}

// Commit pairs packed as state slot | next-state slot << 16, in LDS. With
// more than two wavefronts, wavefronts 0 and 1 take no pairs: they start
// their lane work (inputs, hash) at once, which runs beside the commit.
template <int NT>
__device__ __forceinline__ void bp_commit_lds(uint64_t* L, const lds_u32* c, uint32_t n, uint32_t tid) {
    constexpr uint32_t SKIP = NT > 128 ? 128u : 0u;
    if (tid < SKIP) return;
    for (uint32_t i = tid - SKIP; i < n; i += NT - SKIP) {
        const uint32_t e = c[i];
        L[e & 0xffffu] = L[e >> 16];
    }
}

// Rounds [i0, i1) of the register-resident gate records, a barrier where
// the round table says a level ends.
template <int MAXR, bool BA>
__device__ __forceinline__ void bp_rounds(uint64_t* L, const uint2 (&d)[MAXR], const rf_bp_params& P) {
    // the round count and barrier mask pass an opaque step each cycle, so
    // the per-round tests stay scalar compares here instead of precomputed
    // lane masks (which outgrow the scalar registers and spill)
    uint32_t nr = P.nr;
    uint64_t bar = P.bar;
    __asm__ volatile("" : "+s"(nr), "+s"(bar));
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
        if (i < (int)nr) {
            if constexpr (BA) {
                bp_gate_bytes_live(d[i]);
            } else {
                bp_gate_slots_live(d[i]);
            }
            if ((bar >> i) & 1ull) __syncthreads();
        }
    }
}

// Workgroup of NT threads (NT >= 128) per 64 lanes. Lane t of the group is
// thread t of every wavefront: wavefront 0 generates stimuli, wavefront 1
// hashes outputs, wavefront (2 + i) mod NW serves memory i. After program M
// of cycle n, wavefront 1 hashes cycle n while wavefront 0 generates cycle
// n + 2's inputs.
// MAXR > 0: every thread keeps its gate records of program M (at most
// MAXR rounds) in registers for the whole launch, so the cycle loop issues
// no global-memory loads; MAXR = 0: records are streamed from memory.
// Up to 24 register-resident rounds the kernel fits in 80 registers, which
// leaves room for 6 wavefronts per SIMD (3 workgroups of 512 threads per CU
// when the LDS allows). A 1,024-thread workgroup needs 4 wavefronts per SIMD
// (at most 128 registers each), which holds up to 40 rounds.
template <int NT, int MAXR>
constexpr int bp_waves_per_eu() { return NT >= 1024 ? 4 : (MAXR > 0 && MAXR <= 24 ? 6 : 1); }

template <int NT, int MAXR, bool BA, bool IO>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(bp_waves_per_eu<NT, MAXR>())))
rf_bp_kernel(rf_bp_params P) {
    extern __shared__ uint64_t L[];
    const rf_args& a = P.a;
    const uint32_t tid = threadIdx.x;
    const uint32_t t = tid & 63u;
    const bool w0 = tid < 64u;
    const bool w1 = tid >= 64u && tid < 128u;
    // the 64-lane group this workgroup serves and its cycles this launch
    const bool rot = P.grp_count != 0u;
    const uint32_t grp = rot ? (blockIdx.x + P.grp_base) % P.grp_count : blockIdx.x;
    const uint32_t ncyc = (rot && blockIdx.x >= P.grp_split) ? P.ncyc_hi : a.ncycles;
    const uint32_t lane = grp * 64u + t;
    const bool inb = lane < a.nlanes;
    const uint32_t* prog = P.prog;
    const bool act = inb && a.done[lane] == 0;
    uint64_t h = 0, sid = 0;
    uint32_t cyc = 0;
    if ((w0 || w1) && inb) {
        h = a.hash[lane];
        cyc = a.cyc[lane];
        sid = a.stim[lane];
    }
    constexpr uint32_t NW = NT / 64;
    const uint32_t wave = tid >> 6;
    // cycles this lane advances: ncyc, or with per-stimulus budgets
    // (a.budget) what is left of its budget; the group runs as many cycles
    // as its longest lane needs (every wavefront reduces the same 64 lanes,
    // so no LDS word is needed and the planes start at LDS offset 0)
    const bool bud = a.budget != nullptr;
    uint32_t lim = ncyc;
    uint32_t ngroup = ncyc;
    if (bud) {
        const uint32_t c0 = inb ? a.cyc[lane] : 0u;
        const uint32_t b = act ? a.budget[a.stim[lane]] : 0u;
        lim = b > c0 ? min(ncyc, b - c0) : 0u;
        ngroup = lim;
        for (uint32_t k = 32; k; k >>= 1) ngroup = max(ngroup, (uint32_t)__shfl_xor((int)ngroup, (int)k, 64));
    }
    if (tid == 0) L[0] = 0;
    const uint32_t* st = prog + P.off_state;
    // state planes from the lane-minor arrays, spread over all threads:
    // single-bit variables four threads each (the lanes that exist: lane t
    // of the group is thread t of every wavefront), wider ones by transposes
    {
        bp_load_bits<NT>(L, a, st, P.n_state1, grp, __ballot((int)inb), tid);
        for (uint32_t k2 = P.n_state1 + wave; k2 < P.n_state; k2 += NW) {
            const uint32_t* sl = st + 6u * k2;
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
    const uint32_t* in_tab = prog + P.off_in;
    if (w0 && ncyc) bp_inputs<IO>(L, in_tab, P.n_in, a, sid, cyc, t, act && lim > 0);
    // per-cycle tables into LDS after the planes
    const lds_u32* tc = reinterpret_cast<const lds_u32*>((size_t)P.tab_off);
    const lds_u32* th = tc + P.n_commit[1];
    const lds_u32* ti = th + P.hash_words;
    if (P.tab_off) {
        lds_u32* w = reinterpret_cast<lds_u32*>((size_t)P.tab_off);
        const uint32_t* cm = prog + P.off_commit[1];
        for (uint32_t i = tid; i < P.n_commit[1]; i += NT) w[i] = cm[2 * i] | (cm[2 * i + 1] << 16);
        w += P.n_commit[1];
        const uint32_t* hp = prog + P.off_hash[0];
        for (uint32_t i = tid; i < P.hash_words; i += NT) w[i] = hp[i];
        w += P.hash_words;
        for (uint32_t i = tid; i < 4u * P.n_in; i += NT) w[i] = in_tab[i];
    }
    __syncthreads();
    const uint32_t* lev[3];
    const uint2* gates[3];
    for (int k = 0; k < 3; ++k) {
        lev[k] = prog + P.off_lev[k];
        gates[k] = reinterpret_cast<const uint2*>(prog + P.off_gates[k]);
    }
    constexpr int MR = MAXR > 0 ? MAXR : 1;
    uint2 d[MR];
    if constexpr (MAXR > 0) {
        const uint2* pg = reinterpret_cast<const uint2*>(prog);
#pragma unroll
        for (int i = 0; i < MAXR; ++i) {
            d[i] = make_uint2(0xffffffffu, 0u);
            if (i < (int)P.nr) {
                const uint32_t idx = P.rb[i] + tid;
                if (idx < P.re[i]) d[i] = pg[idx];
            }
        }
    }
    const uint32_t N = ngroup;
    if (N) {
        // cycle 0's next state
        // (the commit writes state slots only; the lane work touches input
        // and output-copy slots only: one barrier covers both)
        bp_phase<NT, BA>(L, lev[RF_BP_A0], gates[RF_BP_A0], P.nlev[RF_BP_A0], tid);
        bp_commit<NT>(L, prog + P.off_commit[0], P.n_commit[0], tid);
        if (P.n_mem) bp_mem_ports(L, P, 0, wave, NW, t, lane, inb, act);
        if (w0) {
            bp_copy_inputs(L, in_tab, P.n_in, t);
            if (N > 1) bp_inputs<IO>(L, in_tab, P.n_in, a, sid, cyc + 1, t, act && lim > 1);
        }
        __syncthreads();
        // cycle n's outputs and cycle n + 1's next state in one pass
        const bool bal = P.progress != nullptr && P.bal_off != 0u;
        lds_u32* balw = reinterpret_cast<lds_u32*>((size_t)P.bal_off);
        for (uint32_t n = 0; n + 1 < N; ++n) {
            if (bal && (n & 7u) == 0u && n) {
                // behind the launch's average: issue ahead of the CU's
                // other workgroups (the oldest workgroup wins otherwise)
                if (*balw) __builtin_amdgcn_s_setprio(3);
                else __builtin_amdgcn_s_setprio(0);
            }
            if constexpr (MAXR > 0)
                bp_rounds<MR, BA>(L, d, P);
            else
                bp_phase<NT, BA>(L, lev[RF_BP_M], gates[RF_BP_M], P.nlev[RF_BP_M], tid);
            if (P.tab_off)
                bp_commit_lds<NT>(L, tc, P.n_commit[1], tid);
            else
                bp_commit<NT>(L, prog + P.off_commit[1], P.n_commit[1], tid);
            if (P.n_mem) bp_mem_ports(L, P, 1, wave, NW, t, lane, inb, act);
            if (w1) {
                const bool cap = IO && a.out.data && act && n < lim;
                const uint64_t hn = P.tab_off ? bp_hash<IO>(L, th, P.n_hash, h, t, a, cap, sid, cyc + n)
                                              : bp_hash<IO>(L, prog + P.off_hash[0], P.n_hash, h, t, a, cap, sid,
                                                            cyc + n);
                if (act && n < lim) h = hn;
            } else if (w0) {
                const bool tick = bal && (n & 7u) == 7u && t == 0u;
                unsigned long long done_all = 0;
                if (tick) done_all = atomicAdd(P.progress, 8ull) + 8ull;
                if (P.tab_off) {
                    bp_copy_inputs(L, ti, P.n_in, t);
                    if (n + 2 < N) bp_inputs<IO>(L, ti, P.n_in, a, sid, cyc + n + 2, t, act && n + 2 < lim);
                } else {
                    bp_copy_inputs(L, in_tab, P.n_in, t);
                    if (n + 2 < N) bp_inputs<IO>(L, in_tab, P.n_in, a, sid, cyc + n + 2, t, act && n + 2 < lim);
                }
                if (tick) *balw = (unsigned long long)(n + 1u) * gridDim.x < done_all ? 1u : 0u;
            }
            __syncthreads();
        }
        if (bal) __builtin_amdgcn_s_setprio(0);
        // the last cycle's outputs
        bp_phase<NT, BA>(L, lev[RF_BP_B], gates[RF_BP_B], P.nlev[RF_BP_B], tid);
        if (w1) {
            const bool cap = IO && a.out.data && act && N - 1 < lim;
            const uint64_t hn = bp_hash<IO>(L, prog + P.off_hash[1], P.n_hash, h, t, a, cap, sid, cyc + N - 1);
            if (act && N - 1 < lim) h = hn;
        }
        __syncthreads();
    }
    // settled state after the last cycle, written back to the lane-minor
    // arrays chunk by chunk (a chunk's values stay in LDS until written);
    // the write-back is spread over all wavefronts: single-bit variables by
    // one broadcast read each, wider ones by transposes (every lane of a
    // wavefront takes part; only active lanes store). With budgets the lanes
    // stopped at different cycles: no state is written back (the hashes are
    // the result).
    const uint32_t* fc = prog + P.off_f;
    // with group rotation, a group's runs before its last write back only
    // the state (the next run loads nothing else)
    const bool full_wb = !rot || (blockIdx.x < P.grp_split ? P.fwb_lo : P.fwb_hi);
    const uint32_t nf = (bud || !full_wb) ? 0u : P.n_f;
    if (!bud && !full_wb) {
        bp_store_bits<NT>(L, a, st, P.n_state1, grp, __ballot((int)act), tid);
        for (uint32_t k2 = P.n_state1 + wave; k2 < P.n_state; k2 += NW) {
            const uint32_t* sl = st + 6u * k2;
            const uint32_t cls = sl[0], words = sl[3], width = sl[4], s0 = sl[5];
            const uint64_t loc = (uint64_t)sl[1] | ((uint64_t)sl[2] << 32);
            if (width <= 64) {
                const uint64_t v = bp_transpose64(t < width ? L[s0 + t] : 0ull, t);
                if (act) bp_store(a, cls, loc, lane, v);
            } else {
                for (uint32_t w = 0; w < words; ++w) {
                    const uint32_t nb = min(32u, width - 32u * w);
                    const uint64_t v = bp_transpose64(t < nb ? L[s0 + 32u * w + t] : 0ull, t);
                    if (act) a.i[(size_t)(loc + w) * a.npad + lane] = (uint32_t)v;
                }
            }
        }
    }
    for (uint32_t c = 0; c < nf; ++c) {
        const uint32_t fnl = fc[0], fng = fc[1], nwb = fc[2], nwb1 = fc[3];
        const uint32_t* flev = fc + RF_BP_FCHUNK_HDR;
        uint32_t go = RF_BP_FCHUNK_HDR + fnl + 1;
        go += go & 1u;
        bp_phase<NT, BA>(L, flev, reinterpret_cast<const uint2*>(fc + go), fnl, tid);
        const uint32_t* wp = fc + go + 2u * fng;
        // single-bit entries (class C), four threads each
        bp_store_bits<NT>(L, a, wp, nwb1, grp, __ballot((int)act), tid);
        wp += 6u * nwb1;
        for (uint32_t k = nwb1; k < nwb; ++k) {
            const uint32_t cls = wp[0], words = wp[3], width = wp[4];
            if ((k - nwb1) % NW == wave) {
                const uint64_t loc = (uint64_t)wp[1] | ((uint64_t)wp[2] << 32);
                const uint32_t* refs = wp + 5;
                if (width <= 64) {
                    const uint64_t v = bp_gather(L, refs, width, t);
                    if (act) bp_store(a, cls, loc, lane, v);
                } else {
                    for (uint32_t w = 0; w < words; ++w) {
                        const uint32_t nb = min(32u, width - 32u * w);
                        const uint64_t v = bp_gather(L, refs + 32u * w, nb, t);
                        if (act) a.i[(size_t)(loc + w) * a.npad + lane] = (uint32_t)v;
                    }
                }
            }
            wp += 5 + width;
        }
        __syncthreads();
        fc += fc[4];
    }
    if (w1 && act) {
        a.hash[lane] = h;
        a.cyc[lane] = cyc + lim;
        if (bud) a.out_hash[sid] = h;
    }
}

}  // namespace

static bool bp_check_phase(const uint32_t* lv, uint32_t nlev, const uint32_t* g, uint32_t ngates,
                           uint32_t nslots, bool byte_addr) {
    if (lv[0] != 0 || lv[nlev] != ngates) return false;
    for (uint32_t l = 0; l < nlev; ++l)
        if (lv[l + 1] < lv[l]) return false;
    for (uint32_t i = 0; i < ngates; ++i) {
        const uint64_t w = (uint64_t)g[2 * i] | ((uint64_t)g[2 * i + 1] << 32);
        if (byte_addr) {
            // dst and s: multiples of 8; x and y: multiples of 8 plus the flag bit
            const uint32_t f[4] = {(uint32_t)(w & 0xffff), (uint32_t)((w >> 16) & 0xffff),
                                   (uint32_t)((w >> 32) & 0xfffe), (uint32_t)((w >> 48) & 0xfffe)};
            for (uint32_t x : f)
                if ((x & 7u) || (x >> 3) >= nslots) return false;
        } else {
            const uint32_t f[4] = {(uint32_t)(w & 0xffff), (uint32_t)((w >> 16) & 0xffff),
                                   (uint32_t)((w >> 32) & 0x7fff), (uint32_t)((w >> 48) & 0x7fff)};
            for (uint32_t x : f)
                if (x >= nslots) return false;
        }
    }
    return true;
}

// A write-back / hash value width and byte count that the kernel handles.
static bool bp_check_refs(const uint32_t* refs, uint32_t width, uint32_t nslots) {
    for (uint32_t i = 0; i < width; ++i)
        if ((refs[i] & 0x7fffffffu) >= nslots) return false;
    return true;
}

int rf_bp_parse(const uint32_t* b, uint64_t words, rf_bp_params* p, const char** why) {
#define BP_FAIL(msg)   \
    do {               \
        *why = (msg);  \
        return -1;     \
    } while (0)
    if (words < RF_BP_HDR_WORDS || b[0] != RF_BP_MAGIC || (b[1] & 0xffffu) != RF_BP_VERSION ||
        (b[1] >> 16) > 1u)
        BP_FAIL("bad bit-plane program header");
    p->byte_addr = b[1] >> 16;
    p->nslots = b[2];
    p->n_in = b[3];
    p->n_state = b[4];
    p->n_commit[0] = b[5];
    p->n_commit[1] = b[16];
    p->n_hash = b[6];
    p->n_f = b[7];
    uint32_t ngates[3];
    for (int k = 0; k < 3; ++k) {
        p->nlev[k] = b[8 + k];
        ngates[k] = b[11 + k];
    }
    p->n_state1 = b[14];
    if (b[15] != words) BP_FAIL("bit-plane program size mismatch");
    if (p->n_state1 > p->n_state) BP_FAIL("bit-plane single-bit count exceeds the state list");
    if (p->nslots == 0 || p->nslots > 0x7fffu || (uint64_t)p->nslots * 8u > 160u * 1024u - 64u)
        BP_FAIL("bit-plane program needs more LDS than a workgroup has");
    if (p->byte_addr && p->nslots > 8192u) BP_FAIL("bit-plane byte-form records need at most 8,192 slots");
    auto ev = [](uint64_t x) { return x + (x & 1); };
    uint64_t o = ev(RF_BP_HDR_WORDS);
    p->off_in = (uint32_t)o;
    o += 4ull * p->n_in;
    o = ev(o);
    p->off_state = (uint32_t)o;
    o += 6ull * p->n_state;
    for (int k = 0; k < 2; ++k) {
        o = ev(o);
        p->off_commit[k] = (uint32_t)o;
        o += 2ull * p->n_commit[k];
    }
    for (int k = 0; k < 3; ++k) {
        o = ev(o);
        p->off_lev[k] = (uint32_t)o;
        o += p->nlev[k] + 1ull;
        o = ev(o);
        p->off_gates[k] = (uint32_t)o;
        o += 2ull * ngates[k];
    }
    if (o > words) BP_FAIL("bit-plane sections exceed the program");
    for (int k = 0; k < 3; ++k)
        if (!bp_check_phase(b + p->off_lev[k], p->nlev[k], b + p->off_gates[k], ngates[k], p->nslots,
                            p->byte_addr))
            BP_FAIL("bit-plane program A0/M/B malformed");
    for (int which = 0; which < 2; ++which) {
        o = ev(o);
        p->off_hash[which] = (uint32_t)o;
        for (uint32_t k = 0; k < p->n_hash; ++k) {
            if (o + 2 > words) BP_FAIL("bit-plane hash section truncated");
            const uint32_t nbytes = b[o], width = b[o + 1];
            if (width == 0 || o + 2 + width > words) BP_FAIL("bit-plane hash entry malformed");
            if ((width <= 64 && (nbytes > 8 || nbytes * 8u < width)) ||
                (width > 64 && (nbytes % 4u || nbytes * 8u < width || nbytes * 8u >= width + 32u)))
                BP_FAIL("bit-plane hash entry malformed");
            if (!bp_check_refs(b + o + 2, width, p->nslots)) BP_FAIL("bit-plane hash slot out of range");
            o += 2ull + width;
        }
        if (which == 0) p->hash_words = (uint32_t)(o - p->off_hash[0]);
    }
    o = ev(o);
    p->n_mem = b[17];
    p->off_mem = (uint32_t)o;
    o += (uint64_t)RF_BP_MEM_WORDS * p->n_mem;
    if (o > words) BP_FAIL("bit-plane memory table truncated");
    for (uint32_t k = 0; k < p->n_mem; ++k) {
        const uint32_t* me = b + p->off_mem + RF_BP_MEM_WORDS * k;
        if (me[1] > 3 || me[2] > 1 || me[6] == 0 || me[7] == 0 || me[7] > 64 ||
            me[7] > 8u * (1u << me[1]) || (me[2] && (uint64_t)me[6] << me[1] > me[5]))
            BP_FAIL("bit-plane memory entry malformed");
    }
    for (int which = 0; which < 2; ++which) {
        o = ev(o);
        p->off_ports[which] = (uint32_t)o;
        if (o + 2 > words || b[o] < 2 || o + b[o] > words) BP_FAIL("bit-plane port list malformed");
        const uint64_t end = o + b[o];
        uint64_t q = o + 2;
        for (uint32_t r = 0; r < b[o + 1]; ++r) {
            if (q + 5 > end) BP_FAIL("bit-plane port list truncated");
            const uint32_t kind = b[q], mi = b[q + 1], na = b[q + 2], n2 = b[q + 3];
            const uint64_t len = (kind == 0 ? 4u : 5u) + (uint64_t)na + n2;
            if (kind > 1 || mi >= p->n_mem || na == 0 || na > 64 || n2 > 64 || q + len > end)
                BP_FAIL("bit-plane port malformed");
            if (kind == 0) {
                if (!bp_check_refs(b + q + 4, na, p->nslots)) BP_FAIL("bit-plane port slot out of range");
                for (uint32_t j = 0; j < n2; ++j) {
                    const uint32_t tg = b[q + 4 + na + j];
                    if ((tg >> 24) >= b[p->off_mem + RF_BP_MEM_WORDS * mi + 7] || (tg & 0xffffffu) >= p->nslots)
                        BP_FAIL("bit-plane read port target out of range");
                }
            } else {
                if (n2 != b[p->off_mem + RF_BP_MEM_WORDS * mi + 7] || !bp_check_refs(b + q + 4, 1 + na + n2, p->nslots))
                    BP_FAIL("bit-plane write port malformed");
            }
            q += len;
        }
        if (q != end) BP_FAIL("bit-plane port list size mismatch");
        o = end;
    }
    o = ev(o);
    p->off_f = (uint32_t)o;
    for (uint32_t c = 0; c < p->n_f; ++c) {
        if (o + RF_BP_FCHUNK_HDR > words) BP_FAIL("bit-plane write-back chunk truncated");
        const uint32_t nl = b[o], ng = b[o + 1], nwb = b[o + 2], nwb1 = b[o + 3], cw = b[o + 4];
        if (nwb1 > nwb || cw < RF_BP_FCHUNK_HDR || o + cw > words) BP_FAIL("bit-plane write-back chunk malformed");
        uint64_t g = o + RF_BP_FCHUNK_HDR + nl + 1;
        g = ev(g);
        uint64_t e = g + 2ull * ng;
        if (e > o + cw) BP_FAIL("bit-plane write-back chunk malformed");
        if (!bp_check_phase(b + o + RF_BP_FCHUNK_HDR, nl, b + g, ng, p->nslots, p->byte_addr))
            BP_FAIL("bit-plane write-back phase malformed");
        for (uint32_t k = 0; k < nwb; ++k) {
            if (e + 5 > o + cw) BP_FAIL("bit-plane write-back entry truncated");
            const uint32_t cls = b[e], wds = b[e + 3], width = b[e + 4];
            if ((k < nwb1) != (width == 1)) BP_FAIL("bit-plane write-back list not ordered single-bit first");
            if (width == 1 && cls != RF_CLS_C) BP_FAIL("bit-plane single-bit write-back entry not of class C");
            if (cls > 3 || width == 0 || e + 5 + width > o + cw ||
                (width > 64 && (cls != RF_CLS_I || wds * 32u < width || wds * 32u >= width + 32u)))
                BP_FAIL("bit-plane write-back entry malformed");
            if (!bp_check_refs(b + e + 5, width, p->nslots)) BP_FAIL("bit-plane write-back slot out of range");
            e += 5ull + width;
        }
        o += cw;
    }
    if (ev(o) != words) BP_FAIL("bit-plane program sections do not add up");
    for (uint32_t k = 0; k < p->n_in; ++k) {
        const uint32_t* ic = b + p->off_in + 4 * k;
        if (ic[1] == 0 || ic[1] > 64 || ic[2] + ic[1] > p->nslots || ic[3] + ic[1] > p->nslots)
            BP_FAIL("bit-plane input slot out of range");
    }
    for (uint32_t k = 0; k < p->n_state; ++k) {
        const uint32_t* sl = b + p->off_state + 6 * k;
        if ((k < p->n_state1) != (sl[4] == 1)) BP_FAIL("bit-plane state list not ordered single-bit first");
        if (sl[4] == 1 && sl[0] != RF_CLS_C) BP_FAIL("bit-plane single-bit state entry not of class C");
        if (sl[0] > 3 || sl[4] == 0 || sl[5] + sl[4] > p->nslots ||
            (sl[4] > 64 && (sl[0] != RF_CLS_I || sl[3] * 32u < sl[4] || sl[3] * 32u >= sl[4] + 32u)))
            BP_FAIL("bit-plane state entry out of range");
    }
    for (int which = 0; which < 2; ++which)
        for (uint32_t k = 0; k < p->n_commit[which]; ++k) {
            const uint32_t* c = b + p->off_commit[which] + 2 * k;
            if (c[0] >= p->nslots || c[1] >= p->nslots) BP_FAIL("bit-plane commit slot out of range");
        }
    // the per-cycle tables go to LDS after the planes when that costs no
    // workgroup per CU (160 KB of LDS per CU)
    const uint64_t planes = (uint64_t)p->nslots * 8u;
    const uint64_t tab = 4ull * (p->n_commit[1] + (uint64_t)p->hash_words + 4ull * p->n_in);
    const uint64_t cu = 160u * 1024u;
    p->tab_off = (planes + tab + 8u <= cu - 64u && cu / planes == cu / (planes + tab + 8u)) ? (uint32_t)planes : 0u;
    // the progress-balancing word after them, when it costs no workgroup per CU
    const uint64_t used = planes + (p->tab_off ? tab : 0u);
    const uint64_t bal = (used + 7u) & ~7ull;
    p->bal_off = (bal + 8u <= cu - 64u && cu / planes == cu / (bal + 8u)) ? (uint32_t)bal : 0u;
    return 0;
#undef BP_FAIL
}

void rf_bp_rounds(const uint32_t* b, uint32_t threads, rf_bp_params* p) {
    p->nr = 0;
    p->bar = 0;
    uint32_t r = 0;
    const uint32_t* lv = b + p->off_lev[RF_BP_M];
    const uint32_t g = p->off_gates[RF_BP_M] / 2;   // uint2 index of the first record
    for (uint32_t l = 0; l < p->nlev[RF_BP_M]; ++l) {
        const uint32_t lo = lv[l], hi = lv[l + 1];
        for (uint32_t s = lo; s < hi; s += threads) {
            if (r >= RF_BP_MAXR) {
                p->bar = 0;
                return;
            }
            p->rb[r] = g + s;
            p->re[r] = g + hi;
            if (s + threads >= hi) p->bar |= 1ull << r;
            ++r;
        }
    }
    p->nr = r;
}

static uint32_t bp_tab_words(const rf_bp_params& p) {
    return p.n_commit[1] + p.hash_words + 4u * p.n_in;
}

static size_t bp_shmem(const rf_bp_params& p) {
    const size_t base = (size_t)p.nslots * 8u + (p.tab_off ? 4u * (size_t)bp_tab_words(p) : 0u);
    return p.bal_off ? std::max(base, (size_t)p.bal_off + 8u) : base;
}

// Launches the instance, or with ``occ`` set only reports how many of its
// workgroups one CU holds at once.
template <int NT, int MAXR, bool BA, bool IO>
static hipError_t bp_launch_io(const rf_bp_params& p, hipStream_t stream, int* occ) {
    const size_t shmem = bp_shmem(p);
    const void* fn = reinterpret_cast<const void*>(&rf_bp_kernel<NT, MAXR, BA, IO>);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 64);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, fn, NT, shmem);
    const uint32_t grid = p.grp_count ? p.grp_n : (p.a.nlanes + 63u) / 64u;
    hipLaunchKernelGGL((rf_bp_kernel<NT, MAXR, BA, IO>), dim3(grid), dim3(NT), shmem, stream, p);
    return hipGetLastError();
}

// The stream-capable instance only for launches with an input stream or an
// output capture: its extra registers cost the generator-only launches
// spills in the cycle loop.
template <int NT, int MAXR, bool BA>
static hipError_t bp_launch_ba(const rf_bp_params& p, hipStream_t stream, int* occ) {
    return (p.a.in.data || p.a.out.data) ? bp_launch_io<NT, MAXR, BA, true>(p, stream, occ)
                                         : bp_launch_io<NT, MAXR, BA, false>(p, stream, occ);
}

template <int NT, int MAXR>
static hipError_t bp_launch_k(const rf_bp_params& p, hipStream_t stream, int* occ) {
    return p.byte_addr ? bp_launch_ba<NT, MAXR, true>(p, stream, occ)
                       : bp_launch_ba<NT, MAXR, false>(p, stream, occ);
}

template <int NT>
static hipError_t bp_launch_nt(const rf_bp_params& p, hipStream_t stream, int* occ) {
    const uint32_t r = p.nr;
    if (r == 0 || r > RF_BP_MAXR || (NT >= 1024 && r > 40)) return bp_launch_k<NT, 0>(p, stream, occ);
    if (r <= 16) return bp_launch_k<NT, 16>(p, stream, occ);
    if (r <= 24) return bp_launch_k<NT, 24>(p, stream, occ);
    if (r <= 32) return bp_launch_k<NT, 32>(p, stream, occ);
    if (NT >= 1024 || r <= 40) return bp_launch_k<NT, 40>(p, stream, occ);
    return bp_launch_k<NT, (NT >= 1024 ? 40 : 48)>(p, stream, occ);
}

static hipError_t bp_dispatch(const rf_bp_params& p, uint32_t threads, hipStream_t stream, int* occ) {
    switch (threads) {
        case 128: return bp_launch_k<128, 0>(p, stream, occ);
        case 256: return bp_launch_nt<256>(p, stream, occ);
        case 512: return bp_launch_nt<512>(p, stream, occ);
        case 1024: return bp_launch_nt<1024>(p, stream, occ);
        default: return hipErrorInvalidValue;
    }
}

hipError_t rf_bp_launch(const rf_bp_params& p, uint32_t threads, hipStream_t stream) {
    if (p.grp_count && (p.grp_n == 0 || p.grp_n > p.grp_count || p.grp_base >= p.grp_count ||
                        p.grp_count != (p.a.nlanes + 63u) / 64u))
        return hipErrorInvalidValue;
    return bp_dispatch(p, threads, stream, nullptr);
}

hipError_t rf_bp_occupancy(const rf_bp_params& p, uint32_t threads, int* per_cu) {
    if (!per_cu) return hipErrorInvalidValue;
    *per_cu = 0;
    return bp_dispatch(p, threads, nullptr, per_cu);
}
