// rf_ops.h -- word-exact two-state Verilog integer operations for generated
// gfx950 kernels (and the one-lane host build of the same generated code).
//
// Semantics follow the reference's device op library include/rf_verilated.h
// (VL_* helpers): values are clean (bits at or above the width are zero),
// shifts by at least the width give 0 (VL_SHIFTL_WWI :2519-2533, VL_SHIFTR_WWI
// :2590), arithmetic right shifts fill with the sign bit of the operand width
// (VL_SHIFTRS_III :2657-2666), signed compares sign-extend from the operand
// width (VL_GTS_III :1761-1817), wide add/sub propagate carries word by word
// (VL_ADD_W :1864-1874, VL_SUB_W :1876-1889), division by zero yields 0
// (VL_DIV_III :1856). Unlike the reference, every wide helper here is usable
// on the device (the reference leaves VL_SUB_W, VL_MUL_W, VL_EQ_W, VL_AND_W
// and the variable wide shifts host-only, include/rf_verilated.h:1698-1905).
//
// Narrow values (<= 64 bits) are uint32_t (<= 32) or uint64_t; wide values
// (> 64 bits) are arrays of uint32_t words, least significant word first.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RF_HD __host__ __device__ __forceinline__
#else
#define RF_HD static inline
#endif

// An optimisation barrier on a value: after it the compiler cannot reuse
// values derived from the old one. The whole-cycle kernels put one on the
// lane's slot base pointers and stride before their final stores, so the
// slot addresses of the prologue's loads are recomputed there instead of
// being kept live (and spilled) across the cycle loop.
#if defined(__HIP_DEVICE_COMPILE__)
#define RF_OPAQUE_S(x) asm volatile("" : "+s"(x))
#define RF_OPAQUE_V(x) asm volatile("" : "+v"(x))
#else
#define RF_OPAQUE_S(x) asm volatile("" : "+r"(x))
#define RF_OPAQUE_V(x) asm volatile("" : "+r"(x))
#endif

// ---------------------------------------------------------------- masks
RF_HD uint32_t rf_m32(int w) { return w >= 32 ? 0xffffffffu : ((1u << w) - 1u); }
RF_HD uint64_t rf_m64(int w) { return w >= 64 ? ~0ull : ((1ull << w) - 1ull); }

// ------------------------------------------------------------- narrow ops
RF_HD uint32_t rf_shl32(uint32_t a, uint64_t b, int w) {
    return b >= (uint64_t)w ? 0u : ((a << (uint32_t)b) & rf_m32(w));
}
RF_HD uint64_t rf_shl64(uint64_t a, uint64_t b, int w) {
    return b >= (uint64_t)w ? 0ull : ((a << b) & rf_m64(w));
}
RF_HD uint32_t rf_shr32(uint32_t a, uint64_t b, int w) {
    return b >= (uint64_t)w ? 0u : (a >> (uint32_t)b);
}
RF_HD uint64_t rf_shr64(uint64_t a, uint64_t b, int w) {
    return b >= (uint64_t)w ? 0ull : (a >> b);
}
RF_HD int32_t rf_sx32(uint32_t a, int w) {
    return w >= 32 ? (int32_t)a : ((int32_t)(a << (32 - w)) >> (32 - w));
}
RF_HD int64_t rf_sx64(uint64_t a, int w) {
    return w >= 64 ? (int64_t)a : ((int64_t)(a << (64 - w)) >> (64 - w));
}
RF_HD uint32_t rf_sra32(uint32_t a, uint64_t b, int w) {
    const int32_t s = rf_sx32(a, w);
    const uint32_t sh = b > 31u ? 31u : (uint32_t)b;
    return (uint32_t)(s >> sh) & rf_m32(w);
}
RF_HD uint64_t rf_sra64(uint64_t a, uint64_t b, int w) {
    const int64_t s = rf_sx64(a, w);
    const uint64_t sh = b > 63u ? 63u : b;
    return (uint64_t)(s >> sh) & rf_m64(w);
}
RF_HD uint32_t rf_divs32(uint32_t a, uint32_t b, int w) {
    if (b == 0) return 0;
    const int64_t sa = rf_sx32(a, w), sb = rf_sx32(b, w);
    return (uint32_t)(sa / sb) & rf_m32(w);
}
RF_HD uint32_t rf_mods32(uint32_t a, uint32_t b, int w) {
    if (b == 0) return 0;
    const int64_t sa = rf_sx32(a, w), sb = rf_sx32(b, w);
    return (uint32_t)(sa % sb) & rf_m32(w);
}
RF_HD uint64_t rf_divs64(uint64_t a, uint64_t b, int w) {
    if (b == 0) return 0;
    const int64_t sa = rf_sx64(a, w), sb = rf_sx64(b, w);
    if (sb == -1) return (0ull - (uint64_t)sa) & rf_m64(w);  // avoids INT64_MIN / -1
    return (uint64_t)(sa / sb) & rf_m64(w);
}
RF_HD uint64_t rf_mods64(uint64_t a, uint64_t b, int w) {
    if (b == 0) return 0;
    const int64_t sa = rf_sx64(a, w), sb = rf_sx64(b, w);
    if (sb == -1) return 0;
    return (uint64_t)(sa % sb) & rf_m64(w);
}
RF_HD uint32_t rf_parity32(uint32_t a) { return (uint32_t)__builtin_popcount(a) & 1u; }
RF_HD uint32_t rf_parity64(uint64_t a) { return (uint32_t)__builtin_popcountll(a) & 1u; }
// bit slot b of a packed register of 1-bit values (v must be 0 or 1)
RF_HD uint32_t rf_bget(uint32_t w, int b) { return (w >> b) & 1u; }
RF_HD uint32_t rf_bset(uint32_t w, int b, uint32_t v) { return (w & ~(1u << b)) | (v << b); }
RF_HD uint32_t rf_divu32(uint32_t a, uint32_t b) { return b ? a / b : 0u; }
RF_HD uint32_t rf_modu32(uint32_t a, uint32_t b) { return b ? a % b : 0u; }
RF_HD uint64_t rf_divu64(uint64_t a, uint64_t b) { return b ? a / b : 0ull; }
RF_HD uint64_t rf_modu64(uint64_t a, uint64_t b) { return b ? a % b : 0ull; }
// dst (width wd <= 64) with bits [lsb +: wl] replaced by v; bits at or above
// wd are dropped, a select starting at or beyond wd writes nothing
RF_HD uint64_t rf_ins64(uint64_t dst, int wd, uint64_t lsb, int wl, uint64_t v) {
    if (lsb >= (uint64_t)wd) return dst;
    const uint64_t m = rf_m64(wl) << lsb;
    return ((dst & ~m) | ((v << lsb) & m)) & rf_m64(wd);
}

// --------------------------------------------------------------- wide ops
// n = words, w = bit width; outputs are cleaned (top word masked).
RF_HD int rfw_words(int w) { return (w + 31) >> 5; }
RF_HD void rfw_clean(uint32_t* o, int w) { o[rfw_words(w) - 1] &= rf_m32(((w - 1) & 31) + 1); }
RF_HD void rfw_zero(uint32_t* o, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = 0;
}
RF_HD void rfw_copy(uint32_t* o, const uint32_t* a, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = a[i];
}
RF_HD void rfw_from64(uint32_t* o, int w, uint64_t v) {
    o[0] = (uint32_t)v;
    o[1] = (uint32_t)(v >> 32);
    for (int i = 2; i < rfw_words(w); ++i) o[i] = 0;
}
// zero / sign extension of a narrow (<= 64-bit, width wa) value to wide w
RF_HD void rfw_sext64(uint32_t* o, int w, uint64_t v, int wa) {
    const int64_t s = rf_sx64(v, wa);
    const uint32_t fill = s < 0 ? 0xffffffffu : 0u;
    o[0] = (uint32_t)(uint64_t)s;
    o[1] = (uint32_t)((uint64_t)s >> 32);
    for (int i = 2; i < rfw_words(w); ++i) o[i] = fill;
    rfw_clean(o, w);
}
// wide -> wider
RF_HD void rfw_zext(uint32_t* o, int w, const uint32_t* a, int wa) {
    const int na = rfw_words(wa);
    for (int i = 0; i < rfw_words(w); ++i) o[i] = i < na ? a[i] : 0u;
}
RF_HD void rfw_sextw(uint32_t* o, int w, const uint32_t* a, int wa) {
    const int na = rfw_words(wa);
    const uint32_t sign = (a[(wa - 1) >> 5] >> ((wa - 1) & 31)) & 1u;
    const uint32_t fill = sign ? 0xffffffffu : 0u;
    for (int i = 0; i < rfw_words(w); ++i) o[i] = i < na ? a[i] : fill;
    if (sign && (wa & 31)) o[na - 1] |= ~rf_m32(wa & 31);
    rfw_clean(o, w);
}
RF_HD void rfw_and(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = a[i] & b[i];
}
RF_HD void rfw_or(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = a[i] | b[i];
}
RF_HD void rfw_xor(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = a[i] ^ b[i];
}
RF_HD void rfw_xnor(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = ~(a[i] ^ b[i]);
    rfw_clean(o, w);
}
RF_HD void rfw_not(uint32_t* o, const uint32_t* a, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = ~a[i];
    rfw_clean(o, w);
}
RF_HD void rfw_add(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    uint64_t carry = 0;
    for (int i = 0; i < rfw_words(w); ++i) {
        carry += (uint64_t)a[i] + (uint64_t)b[i];
        o[i] = (uint32_t)carry;
        carry >>= 32;
    }
    rfw_clean(o, w);
}
RF_HD void rfw_sub(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    uint64_t carry = 1;
    for (int i = 0; i < rfw_words(w); ++i) {
        carry += (uint64_t)a[i] + (uint64_t)(uint32_t)~b[i];
        o[i] = (uint32_t)carry;
        carry >>= 32;
    }
    rfw_clean(o, w);
}
RF_HD void rfw_neg(uint32_t* o, const uint32_t* a, int w) {
    uint64_t carry = 1;
    for (int i = 0; i < rfw_words(w); ++i) {
        carry += (uint64_t)(uint32_t)~a[i];
        o[i] = (uint32_t)carry;
        carry >>= 32;
    }
    rfw_clean(o, w);
}
// o must not alias a or b
RF_HD void rfw_mul(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    const int n = rfw_words(w);
    for (int i = 0; i < n; ++i) o[i] = 0;
    for (int i = 0; i < n; ++i) {
        uint64_t carry = 0;
        for (int j = 0; i + j < n; ++j) {
            const uint64_t t = (uint64_t)a[i] * (uint64_t)b[j] + (uint64_t)o[i + j] + carry;
            o[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    rfw_clean(o, w);
}
// o must not alias a
RF_HD void rfw_shl(uint32_t* o, const uint32_t* a, uint64_t sh, int w) {
    const int n = rfw_words(w);
    if (sh >= (uint64_t)w) { rfw_zero(o, w); return; }
    const int ws = (int)(sh >> 5), bs = (int)(sh & 31);
    for (int i = n - 1; i >= 0; --i) {
        const int s = i - ws;
        uint32_t v = 0;
        if (s >= 0) v = a[s] << bs;
        if (bs && s - 1 >= 0) v |= a[s - 1] >> (32 - bs);
        o[i] = v;
    }
    rfw_clean(o, w);
}
RF_HD void rfw_shr(uint32_t* o, const uint32_t* a, uint64_t sh, int w) {
    const int n = rfw_words(w);
    if (sh >= (uint64_t)w) { rfw_zero(o, w); return; }
    const int ws = (int)(sh >> 5), bs = (int)(sh & 31);
    for (int i = 0; i < n; ++i) {
        const int s = i + ws;
        uint32_t v = 0;
        if (s < n) v = a[s] >> bs;
        if (bs && s + 1 < n) v |= a[s + 1] << (32 - bs);
        o[i] = v;
    }
}
RF_HD uint32_t rfw_bit(const uint32_t* a, int b) { return (a[b >> 5] >> (b & 31)) & 1u; }
RF_HD void rfw_sra(uint32_t* o, const uint32_t* a, uint64_t sh, int w) {
    const uint32_t sign = rfw_bit(a, w - 1);
    const uint64_t s = sh >= (uint64_t)w ? (uint64_t)w : sh;
    if (s >= (uint64_t)w) {
        for (int i = 0; i < rfw_words(w); ++i) o[i] = sign ? 0xffffffffu : 0u;
        rfw_clean(o, w);
        return;
    }
    rfw_shr(o, a, s, w);
    if (sign) {
        // set bits [w - s, w)
        for (int b = w - (int)s; b < w; ++b) o[b >> 5] |= 1u << (b & 31);
    }
}
RF_HD uint32_t rfw_eq(const uint32_t* a, const uint32_t* b, int w) {
    uint32_t d = 0;
    for (int i = 0; i < rfw_words(w); ++i) d |= a[i] ^ b[i];
    return d == 0;
}
RF_HD int rfw_cmpu(const uint32_t* a, const uint32_t* b, int w) {
    for (int i = rfw_words(w) - 1; i >= 0; --i) {
        if (a[i] > b[i]) return 1;
        if (a[i] < b[i]) return -1;
    }
    return 0;
}
RF_HD int rfw_cmps(const uint32_t* a, const uint32_t* b, int w) {
    const uint32_t sa = rfw_bit(a, w - 1), sb = rfw_bit(b, w - 1);
    if (sa != sb) return sa ? -1 : 1;
    return rfw_cmpu(a, b, w);
}
RF_HD uint32_t rfw_redor(const uint32_t* a, int w) {
    uint32_t d = 0;
    for (int i = 0; i < rfw_words(w); ++i) d |= a[i];
    return d != 0;
}
RF_HD uint32_t rfw_redand(const uint32_t* a, int w) {
    const int n = rfw_words(w);
    uint32_t d = 0xffffffffu;
    for (int i = 0; i < n - 1; ++i) d &= a[i];
    d &= a[n - 1] | ~rf_m32(((w - 1) & 31) + 1);
    return d == 0xffffffffu;
}
RF_HD uint32_t rfw_redxor(const uint32_t* a, int w) {
    uint32_t d = 0;
    for (int i = 0; i < rfw_words(w); ++i) d ^= a[i];
    return rf_parity32(d);
}
RF_HD void rfw_cond(uint32_t* o, uint32_t c, const uint32_t* a, const uint32_t* b, int w) {
    for (int i = 0; i < rfw_words(w); ++i) o[i] = c ? a[i] : b[i];
}
// bits [lsb, lsb + wr) of a (width wa) as a <= 64-bit value; bits >= wa read 0
RF_HD uint64_t rfw_get64(const uint32_t* a, int wa, uint64_t lsb, int wr) {
    if (lsb >= (uint64_t)wa) return 0;
    const int n = rfw_words(wa);
    const int wi = (int)(lsb >> 5), bi = (int)(lsb & 31);
    const uint64_t lo = a[wi];
    const uint64_t mid = wi + 1 < n ? a[wi + 1] : 0u;
    const uint64_t hi = wi + 2 < n ? a[wi + 2] : 0u;
    uint64_t r = (lo >> bi) | (mid << (32 - bi));
    if (bi) r |= hi << (64 - bi);
    return r & rf_m64(wr);
}
// o[lsb +: wv] = v (clean, wv <= 64 bits); bits falling beyond wo are dropped
RF_HD void rfw_put64(uint32_t* o, int wo, uint64_t lsb, int wv, uint64_t v) {
    uint64_t pos = lsb;
    int rem = wv;
    while (rem > 0 && pos < (uint64_t)wo) {
        const int wi = (int)(pos >> 5), bi = (int)(pos & 31);
        const int take = (32 - bi) < rem ? (32 - bi) : rem;
        const uint32_t m = rf_m32(take) << bi;
        o[wi] = (o[wi] & ~m) | (((uint32_t)v << bi) & m);
        v = take >= 64 ? 0 : (v >> take);
        pos += (uint64_t)take;
        rem -= take;
    }
    rfw_clean(o, wo);
}
// wide source of width wv inserted at lsb
RF_HD void rfw_putw(uint32_t* o, int wo, uint64_t lsb, const uint32_t* v, int wv) {
    for (int k = 0; k < wv; k += 32) {
        const int take = (wv - k) < 32 ? (wv - k) : 32;
        rfw_put64(o, wo, lsb + (uint64_t)k, take, v[k >> 5] & rf_m32(take));
    }
}
// o = {o[w-1-k:0], v} in place (0 < k <= 64, v clean): a shift register
RF_HD void rfw_shl_in(uint32_t* o, int w, int k, uint64_t v) {
    const int n = rfw_words(w), ws = k >> 5, bs = k & 31;
    for (int i = n - 1; i >= 0; --i) {
        const uint32_t hi = i - ws >= 0 ? o[i - ws] : 0u;
        const uint32_t lo = i - ws - 1 >= 0 ? o[i - ws - 1] : 0u;
        o[i] = bs ? (hi << bs) | (lo >> (32 - bs)) : hi;
    }
    o[0] |= (uint32_t)v;
    if (n > 1 && k > 32) o[1] |= (uint32_t)(v >> 32);
    rfw_clean(o, w);
}
// o (width wr) = a[lsb +: wr]
RF_HD void rfw_extract(uint32_t* o, int wr, const uint32_t* a, int wa, uint64_t lsb) {
    for (int k = 0; k < rfw_words(wr); ++k) o[k] = (uint32_t)rfw_get64(a, wa, lsb + 32ull * k, 32);
    rfw_clean(o, wr);
}
// Wide division (restoring, one quotient bit per step), widths up to
// RFW_DIV_MAX_WORDS * 32 bits; x / 0 = x % 0 = 0 (VL_DIV_WWW / VL_MODDIV_WWW,
// include/rf_verilated.h:1856-1861). q or r may be null.
#define RFW_DIV_MAX_WORDS 16
RF_HD void rfw_divmodu(uint32_t* q, uint32_t* r, const uint32_t* a, const uint32_t* b, int w) {
    const int n = rfw_words(w);
    uint32_t rem[RFW_DIV_MAX_WORDS + 1];
    uint32_t quo[RFW_DIV_MAX_WORDS];
    for (int i = 0; i <= n; ++i) rem[i] = 0;
    for (int i = 0; i < n; ++i) quo[i] = 0;
    uint32_t nz = 0;
    for (int i = 0; i < n; ++i) nz |= b[i];
    if (nz) {
        for (int bit = w - 1; bit >= 0; --bit) {
            // rem = (rem << 1) | a[bit]
            for (int i = n; i > 0; --i) rem[i] = (rem[i] << 1) | (rem[i - 1] >> 31);
            rem[0] = (rem[0] << 1) | rfw_bit(a, bit);
            // rem >= b ?
            int ge = 1;
            for (int i = n; i >= 0; --i) {
                const uint32_t bi = i < n ? b[i] : 0u;
                if (rem[i] != bi) { ge = rem[i] > bi; break; }
            }
            if (ge) {
                uint64_t borrow = 0;
                for (int i = 0; i <= n; ++i) {
                    const uint64_t bi = i < n ? b[i] : 0u;
                    const uint64_t t = (uint64_t)rem[i] - bi - borrow;
                    rem[i] = (uint32_t)t;
                    borrow = (t >> 63) & 1u;
                }
                quo[bit >> 5] |= 1u << (bit & 31);
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        if (q) q[i] = quo[i];
        if (r) r[i] = nz ? rem[i] : 0u;
    }
}
RF_HD void rfw_divu(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    rfw_divmodu(o, nullptr, a, b, w);
}
RF_HD void rfw_modu(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    rfw_divmodu(nullptr, o, a, b, w);
}
// signed: truncating division, remainder takes the dividend's sign
RF_HD void rfw_divmods(uint32_t* o, const uint32_t* a, const uint32_t* b, int w, int want_rem) {
    uint32_t ua[RFW_DIV_MAX_WORDS], ub[RFW_DIV_MAX_WORDS], res[RFW_DIV_MAX_WORDS];
    const uint32_t sa = rfw_bit(a, w - 1), sb = rfw_bit(b, w - 1);
    if (sa) rfw_neg(ua, a, w); else rfw_copy(ua, a, w);
    if (sb) rfw_neg(ub, b, w); else rfw_copy(ub, b, w);
    if (want_rem) {
        rfw_divmodu(nullptr, res, ua, ub, w);
        if (sa) rfw_neg(o, res, w); else rfw_copy(o, res, w);
    } else {
        rfw_divmodu(res, nullptr, ua, ub, w);
        if (sa != sb) rfw_neg(o, res, w); else rfw_copy(o, res, w);
    }
}
RF_HD void rfw_divs(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    rfw_divmods(o, a, b, w, 0);
}
RF_HD void rfw_mods(uint32_t* o, const uint32_t* a, const uint32_t* b, int w) {
    rfw_divmods(o, a, b, w, 1);
}
// shift amount given as a wide value: clamp to a 64-bit amount (>= 2^32 is
// "at least the width" for every width this library supports)
RF_HD uint64_t rfw_amount(const uint32_t* a, int w) {
    for (int i = 2; i < rfw_words(w); ++i)
        if (a[i]) return ~0ull;
    return (uint64_t)a[0] | ((uint64_t)(rfw_words(w) > 1 ? a[1] : 0u) << 32);
}
