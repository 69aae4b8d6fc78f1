// rf_bitplane.h -- bit-plane kernel interface inside librf_verilated.so.
//
// A design lowered to single-bit logic carries a program blob (the
// rf_design.bitplane words; layout in rtlflow_amd/bitplane.py, encode()).
// The kernel evaluates it for 64 stimuli per 64-bit plane word: one
// workgroup per 64 lanes, every node value in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_design.h"

#define RF_BP_DEFAULT_THREADS 512
#define RF_BP_MAGIC 0x50424652u
#define RF_BP_VERSION 3u
#define RF_BP_HDR_WORDS 18u

struct rf_bp_params {
    rf_args a;
    const uint32_t* prog;     // device copy of the blob
    uint32_t nslots;
    uint32_t n_in, n_state, n_commit, n_hash, n_wb;
    uint32_t n_state1, n_wb1;   // leading single-bit entries of the state / write-back lists
    uint32_t off_in, off_state, off_commit, off_hash, off_wb;
    uint32_t nlev[3], off_lev[3], off_gates[3];
};

// Parse and check a blob's header; fills every field but ``a`` and ``prog``.
// Returns 0, or -1 with a reason in *why.
int rf_bp_parse(const uint32_t* blob, uint64_t words, rf_bp_params* p, const char** why);

// Launch one kernel over ceil(nlanes / 64) workgroups of ``threads``
// threads (128, 256, 512 or 1024).
hipError_t rf_bp_launch(const rf_bp_params& p, uint32_t threads, hipStream_t stream);
