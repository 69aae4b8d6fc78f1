// rf_stim.h -- stimulus generator and output hash shared by the generated
// kernels (and by the one-lane host build of the same generated code).
//
// Stimulus: a counter-based generator, integer only, identical on host and
// device:  r(seed, s, c, k) = splitmix64(seed ^ (s * 0x9E3779B97F4A7C15)
//                                             ^ (c << 8) ^ k)
// with splitmix64 the published SplitMix64 finalizer (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014).
// Port p (data inputs in declaration order, clocks excluded), 64-bit chunk j
// of cycle c for stimulus s is r(seed, s, c, (p << 3) | j), masked to the
// port width.
//
// Output hash: FNV-1a 64 (offset basis 0xcbf29ce484222325, prime
// 0x100000001b3) over the little-endian bytes of each output port's storage
// words (1/2/4/8 bytes for the u8/u16/u32/u64 classes, 4 per word for wide
// ports), output ports in declaration order, once per clock cycle.
//
// A user input stream (rf_args.in) replaces the generator: word n of the
// cycle's input words (port p's chunk j, ports in the same order) is read
// from the stream instead of r(seed, s, c, (p << 3) | j). The output capture
// (rf_args.out) stores each output port's 64-bit chunks where the hash is
// folded. Reads outside the stream's window give 0 and captures outside the
// capture's window are dropped. A running lane that reads outside the input
// stream gets 0 and raises RF_FLAG_STREAM, which the runtime reports as an
// error.
#pragma once
#include <stdint.h>

#include "rf_design.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RF_HD __host__ __device__ __forceinline__
#else
#ifndef RF_HD
#define RF_HD static inline
#endif
#endif

#define RF_FNV_BASIS 0xcbf29ce484222325ull
#define RF_FNV_PRIME 0x100000001b3ull
#define RF_GOLDEN 0x9E3779B97F4A7C15ull

RF_HD uint64_t rf_splitmix64(uint64_t x) {
    uint64_t z = x + RF_GOLDEN;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
RF_HD uint64_t rf_stim(uint64_t seed, uint64_t s, uint64_t c, uint64_t k) {
    return rf_splitmix64(seed ^ (s * RF_GOLDEN) ^ (c << 8) ^ k);
}
RF_HD uint64_t rf_fnv_bytes(uint64_t h, uint64_t v, int nbytes) {
    for (int i = 0; i < nbytes; ++i) {
        h ^= (v >> (8 * i)) & 0xffu;
        h *= RF_FNV_PRIME;
    }
    return h;
}

// A running lane asked for an input the stream does not hold.
RF_HD void rf_stream_miss(const rf_args& a) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(&a.flags[RF_FLAG_STREAM], 1u);
#else
    a.flags[RF_FLAG_STREAM] |= 1u;
#endif
}

// Input word n (generator key k) of stimulus s in cycle c. ``live``: the
// lane runs this cycle (a miss is then an error the runtime reports).
RF_HD uint64_t rf_input(const rf_args& a, uint64_t s, uint64_t c, uint64_t k, uint32_t n, bool live = true) {
    if (!a.in.data) return rf_stim(a.seed, s, c, k);
    const uint64_t ds = s - a.in.stim0, dc = c - a.in.cycle0;
    if (ds < a.in.nstim && dc < a.in.ncycles) return a.in.data[(dc * a.in.words + n) * a.in.nstim + ds];
    if (live) rf_stream_miss(a);
    return 0;
}

// Output word n of stimulus s in cycle c into the capture, if it covers them.
RF_HD void rf_capture(const rf_args& a, uint64_t s, uint64_t c, uint32_t n, uint64_t v) {
    const uint64_t ds = s - a.out.stim0, dc = c - a.out.cycle0;
    if (a.out.data && ds < a.out.nstim && dc < a.out.ncycles)
        a.out.data[(dc * a.out.words + n) * a.out.nstim + ds] = v;
}
