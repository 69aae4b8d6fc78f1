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
#define RF_BP_VERSION 8u
#define RF_BP_MEM_WORDS 8u
#define RF_BP_FCHUNK_HDR 6u
#define RF_BP_HDR_WORDS 18u

#define RF_BP_MAXR 48   // most gate rounds of the merged per-cycle pass held in registers

// Programs of a blob: A0 (first cycle's next state), M (outputs of cycle n
// merged with the next state of cycle n + 1), B (last cycle's outputs).
enum { RF_BP_A0 = 0, RF_BP_M = 1, RF_BP_B = 2 };

struct rf_bp_params {
    rf_args a;
    // program M as rounds of one gate per thread, for the kernels that keep
    // every thread's gate records in registers across cycles: round i covers
    // the gate records [rb[i], re[i]) (uint2 index from ``prog``), thread t
    // takes record rb[i] + t; bit i of ``bar``: a barrier follows.
    uint32_t nr;
    uint64_t bar;
    uint32_t rb[RF_BP_MAXR], re[RF_BP_MAXR];
    const uint32_t* prog;     // device copy of the blob
    uint32_t nslots;
    uint32_t n_in, n_state, n_commit[2], n_hash, n_f;
    uint32_t n_state1;          // leading single-bit entries of the state list
    uint32_t n_mem;             // memories with ports (8-word entries at off_mem)
    uint32_t byte_addr;         // gate records in the byte form (fields are LDS byte offsets)
    uint32_t off_in, off_state, off_commit[2], off_hash[2], off_f;
    uint32_t off_mem, off_ports[2];   // memory table; port lists after A0 and M
    uint32_t nlev[3], off_lev[3], off_gates[3];
    uint32_t hash_words;        // words of program M's hashed-output table
    // LDS byte offset of the per-cycle tables (program M's commit pairs
    // packed as state | next << 16, its hashed-output table, the input
    // chunks) when the launch holds them after the planes; 0: read from
    // ``prog``. Set when the extra LDS leaves the workgroups per CU unchanged.
    uint32_t tab_off;
    // group rotation: with grp_count != 0 the launch runs grp_n workgroups,
    // workgroup b serving the 64-lane group (grp_base + b) mod grp_count for
    // a.ncycles cycles when b < grp_split, else for ncyc_hi cycles, and
    // writing back every stored variable (program F) when fwb_lo (b <
    // grp_split) or fwb_hi (the others) is set, else only the state; with
    // grp_count == 0 workgroup b serves group b for a.ncycles cycles and
    // writes back everything
    uint32_t grp_base, grp_count, grp_n, grp_split, ncyc_hi;
    uint32_t fwb_lo, fwb_hi;
    // progress balancing: every 8 cycles a workgroup adds the cycles it has
    // run to *progress (zeroed before the launch) and, while it has run
    // fewer than the launch's average, issues at a raised wave priority;
    // the decision goes to all its wavefronts through the LDS word at byte
    // offset bal_off. Off when progress is null or bal_off is 0.
    unsigned long long* progress;
    uint32_t bal_off;
};

// Parse and check a blob (every slot index, level table and section size);
// fills every field but ``a``, ``prog`` and the round table. Returns 0, or -1
// with a reason in *why.
int rf_bp_parse(const uint32_t* blob, uint64_t words, rf_bp_params* p, const char** why);

// Fill the round table of program M for ``threads`` threads per
// workgroup; nr = 0 when it needs more than RF_BP_MAXR rounds.
void rf_bp_rounds(const uint32_t* blob, uint32_t threads, rf_bp_params* p);

// Launch one kernel over ceil(nlanes / 64) workgroups (grp_n with group
// rotation) of ``threads`` threads (128, 256, 512 or 1024).
hipError_t rf_bp_launch(const rf_bp_params& p, uint32_t threads, hipStream_t stream);

// Workgroups of the kernel instance rf_bp_launch would pick for ``p`` that
// one CU holds at once (LDS, registers and threads), in *per_cu.
hipError_t rf_bp_occupancy(const rf_bp_params& p, uint32_t threads, int* per_cu);
