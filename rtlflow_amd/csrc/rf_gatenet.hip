// rf_gatenet.hip -- gate-level netlist engine (include/rf_gatenet.h).
//
// Program blob (u32 words, written by rtlflow_amd/gatenet.py):
//   header[32]: magic 'RFGT', version, n_nodes, n_pi, n_ff, n_ram_bits, n_po,
//     n_rams, ram_depth, ram_width, abits, n_lev_full, n_lev_cone, n_rec_full,
//     n_rec_cone, n_mtasks, and the word offsets of: the full eval's level
//     starts [n_lev_full + 1], the output cone's level starts
//     [n_lev_cone + 1], ff_d[n_ff], po[n_po], the RAM ports
//     [n_rams][abits + 1 + ram_width] (address bits lsb first, write enable,
//     write data lsb first), the full eval's records, the cone's records, the
//     mtask table [n_mtasks][4] = {first record, records, preds, first pred},
//     the pred list, the mtask-ordered records; then n_rec_mt and the total.
//   Records are uint4 {dst | type << 29, in0, in1, in2} (16-byte aligned):
//     type 0 AND, 1 OR, 2 XOR, 3 NAND, 4 MUX (in0 ? in1 : in2).
// Node ids: [0, n_pi) inputs, then flip-flops, RAM read-data bits, gates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <cstdarg>
#include <cstring>
#include <deque>
#include <vector>

#include "rf_gatenet.h"
#include "rf_stim.h"
#include "rf_transpose.h"
#include "rf_verilated.h"

int rf_set_error(int code, const char* fmt, ...);

namespace {

enum {
    H_MAGIC, H_VERSION, H_NODES, H_PI, H_FF, H_RAMBITS, H_PO, H_RAMS, H_DEPTH, H_WIDTH, H_ABITS,
    H_LEV_FULL, H_LEV_CONE, H_REC_FULL, H_REC_CONE, H_MTASKS, H_OFF_LEV_FULL, H_OFF_LEV_CONE,
    H_OFF_FF, H_OFF_PO, H_OFF_RAM, H_OFF_REC_FULL, H_OFF_REC_CONE, H_OFF_MT, H_OFF_PREDS,
    H_OFF_REC_MT, H_REC_MT, H_TOTAL, H_WORDS = 32
};
constexpr uint32_t kMagic = 0x54474652u;  // 'RFGT'
constexpr uint32_t kVersion = 1;
constexpr uint32_t kDstMask = (1u << 29) - 1;

#define GN_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return rf_set_error(RF_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

__device__ __forceinline__ uint8_t gn_op(uint32_t ty, uint8_t a, uint8_t b, uint8_t c) {
    switch (ty) {
        case 0: return a & b;
        case 1: return a | b;
        case 2: return a ^ b;
        case 3: return (uint8_t)((a & b) ^ 1u);
        default: return a ? b : c;
    }
}

__device__ __forceinline__ uint64_t gn_op64(uint32_t ty, uint64_t a, uint64_t b, uint64_t c) {
    switch (ty) {
        case 0: return a & b;
        case 1: return a | b;
        case 2: return a ^ b;
        case 3: return ~(a & b);
        default: return (a & b) | (~a & c);
    }
}

// ------------------------------------------------------------ per-lane form
// X[node * npad + lane]: one byte per node per lane (lane-minor, coalesced
// across a wavefront). ctl[0] = cycle, ctl[2..3] = seed.

// One mtask: its gates in level order for every lane (the records are the
// same for all lanes, so the loads are scalar).
__global__ void __launch_bounds__(256) gn_lane_mtask(const uint4* __restrict__ rec, uint32_t n,
                                                     uint8_t* __restrict__ X, uint64_t npad,
                                                     uint32_t nlanes) {
    const uint32_t lane = blockIdx.x * 256u + threadIdx.x;
    if (lane >= nlanes) return;
    uint8_t* x = X + lane;
    for (uint32_t k = 0; k < n; ++k) {
        const uint4 r = rec[k];
        const uint32_t ty = r.x >> 29;
        const uint8_t a = x[(size_t)r.y * npad], b = x[(size_t)r.z * npad];
        const uint8_t c = ty == 4 ? x[(size_t)r.w * npad] : (uint8_t)0;
        x[(size_t)(r.x & kDstMask) * npad] = gn_op(ty, a, b, c);
    }
}

__global__ void __launch_bounds__(256) gn_lane_stim(uint8_t* __restrict__ X, uint64_t npad,
                                                    uint32_t nlanes, uint32_t n_pi,
                                                    const uint32_t* __restrict__ ctl) {
    const uint32_t lane = blockIdx.x * 256u + threadIdx.x;
    if (lane >= nlanes) return;
    const uint64_t seed = (uint64_t)ctl[2] | ((uint64_t)ctl[3] << 32);
    const uint32_t cyc = ctl[0];
    for (uint32_t j = 0; j * 64u < n_pi; ++j) {
        const uint64_t w = rf_stim(seed, lane, cyc, j);
        for (uint32_t i = 64u * j; i < n_pi && i < 64u * j + 64u; ++i)
            X[(size_t)i * npad + lane] = (uint8_t)((w >> (i - 64u * j)) & 1u);
    }
}

// The rising edge: flip-flops take their D values, each RAM reads its old
// word into the read-data nodes and writes when enabled (one thread per
// lane; a lane's RAMs are its own rows of M).
__global__ void __launch_bounds__(256) gn_lane_edge(const uint32_t* __restrict__ prog, uint8_t* __restrict__ X,
                                                    uint64_t npad, uint32_t nlanes, uint64_t* __restrict__ M) {
    const uint32_t lane = blockIdx.x * 256u + threadIdx.x;
    if (lane >= nlanes) return;
    uint8_t* x = X + lane;
    const uint32_t n_pi = prog[H_PI], n_ff = prog[H_FF];
    const uint32_t* ffd = prog + prog[H_OFF_FF];
    for (uint32_t i = 0; i < n_ff; ++i) x[(size_t)(n_pi + i) * npad] = x[(size_t)ffd[i] * npad];
    const uint32_t nr = prog[H_RAMS], depth = prog[H_DEPTH], width = prog[H_WIDTH], ab = prog[H_ABITS];
    const uint32_t rd0 = n_pi + n_ff;
    const uint64_t wmask = width >= 64 ? ~0ull : ((1ull << width) - 1ull);
    for (uint32_t r = 0; r < nr; ++r) {
        const uint32_t* rp = prog + prog[H_OFF_RAM] + r * (ab + 1u + width);
        uint32_t addr = 0;
        for (uint32_t k = 0; k < ab; ++k) addr |= (uint32_t)x[(size_t)rp[k] * npad] << k;
        const uint32_t we = x[(size_t)rp[ab] * npad];
        uint64_t wd = 0;
        for (uint32_t k = 0; k < width; ++k) wd |= (uint64_t)x[(size_t)rp[ab + 1 + k] * npad] << k;
        uint64_t* mem = M + ((size_t)lane * nr + r) * depth;
        const uint64_t old = addr < depth ? mem[addr] : 0ull;
        if (we && addr < depth) mem[addr] = wd & wmask;
        for (uint32_t k = 0; k < width; ++k) x[(size_t)(rd0 + r * width + k) * npad] = (uint8_t)((old >> k) & 1u);
    }
}

__global__ void __launch_bounds__(256) gn_lane_hash(const uint32_t* __restrict__ prog,
                                                    const uint8_t* __restrict__ X, uint64_t npad,
                                                    uint32_t nlanes, uint64_t* __restrict__ hash,
                                                    const uint32_t* __restrict__ budget,
                                                    const uint32_t* __restrict__ ctl) {
    const uint32_t lane = blockIdx.x * 256u + threadIdx.x;
    if (lane >= nlanes || ctl[0] >= budget[lane]) return;
    const uint32_t n_po = prog[H_PO];
    const uint32_t* po = prog + prog[H_OFF_PO];
    uint64_t v = 0;
    for (uint32_t k = 0; k < n_po; ++k) v |= (uint64_t)X[(size_t)po[k] * npad + lane] << k;
    const int nb = n_po <= 8 ? 1 : n_po <= 16 ? 2 : n_po <= 32 ? 4 : 8;
    hash[lane] = rf_fnv_bytes(hash[lane], v, nb);
}

// ---------------------------------------------------------- bit-plane form
// P[node * G + g]: the plane of 64 lanes (group g) of a node; G groups, a
// power of two, so a wavefront of one gate covers 64 groups with one 512 B
// access per operand.

// A launch covers the first 2^le groups (those with lanes still running;
// the stimuli are dealt to lanes longest budget first): thread t evaluates
// gate t >> le for group t & (2^le - 1); planes keep their 2^lg-group stride.
__global__ void __launch_bounds__(256) gn_bp_gates(const uint4* __restrict__ rec, uint32_t n,
                                                   uint64_t* __restrict__ P, uint32_t lg, uint32_t le) {
    const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint32_t k = (uint32_t)(t >> le);
    if (k >= n) return;
    const uint32_t g = (uint32_t)t & ((1u << le) - 1u);
    const uint4 r = rec[k];
    const uint32_t ty = r.x >> 29;
    const uint64_t a = P[((size_t)r.y << lg) + g], b = P[((size_t)r.z << lg) + g];
    const uint64_t c = ty == 4 ? P[((size_t)r.w << lg) + g] : 0ull;
    P[((size_t)(r.x & kDstMask) << lg) + g] = gn_op64(ty, a, b, c);
}

// One wavefront per group: lane t's primary-input word, transposed into
// planes (thread i ends with input bit i of the group's 64 lanes); a lane
// runs stimulus sid[lane].
__global__ void __launch_bounds__(64) gn_bp_stim(uint64_t* __restrict__ P, uint32_t lg, uint32_t n_pi,
                                                 const uint32_t* __restrict__ ctl,
                                                 const uint32_t* __restrict__ sid, uint32_t nlanes) {
    const uint32_t g = blockIdx.x, t = threadIdx.x;
    const uint64_t seed = (uint64_t)ctl[2] | ((uint64_t)ctl[3] << 32);
    const uint32_t cyc = ctl[0];
    const uint32_t lane = g * 64u + t;
    const uint64_t s = lane < nlanes ? sid[lane] : lane;
    for (uint32_t j = 0; j * 64u < n_pi; ++j) {
        const uint64_t col = bp_transpose64(rf_stim(seed, s, cyc, j), t);
        if (64u * j + t < n_pi) P[((size_t)(64u * j + t) << lg) + g] = col;
    }
}

// The rising edge. Blocks [0, ffb): flip-flops, one thread per (flip-flop,
// group). Blocks [ffb, ffb + G * n_rams): one wavefront per (group, RAM):
// address / enable and data planes transposed to lanes, each lane's RAM row
// read and written, the old words transposed back into read-data planes.
// Flip-flops read gates and write flip-flop nodes; RAM ports read gates and
// write read-data nodes: no block reads what another writes.
__global__ void __launch_bounds__(64) gn_bp_edge(const uint32_t* __restrict__ prog, uint64_t* __restrict__ P,
                                                 uint32_t lg, uint32_t le, uint32_t ffb,
                                                 uint64_t* __restrict__ M) {
    const uint32_t t = threadIdx.x;
    const uint32_t n_pi = prog[H_PI], n_ff = prog[H_FF];
    const uint32_t E = 1u << le;     // groups this launch covers
    if (blockIdx.x < ffb) {
        const uint64_t i = (uint64_t)blockIdx.x * 64u + t;
        const uint32_t f = (uint32_t)(i >> le), g = (uint32_t)i & (E - 1u);
        if (f >= n_ff) return;
        const uint32_t d = prog[prog[H_OFF_FF] + f];
        P[((size_t)(n_pi + f) << lg) + g] = P[((size_t)d << lg) + g];
        return;
    }
    const uint32_t b = blockIdx.x - ffb;
    const uint32_t g = b & (E - 1u), r = b >> le;
    const uint32_t nr = prog[H_RAMS], depth = prog[H_DEPTH], width = prog[H_WIDTH], ab = prog[H_ABITS];
    if (r >= nr) return;
    const uint32_t* rp = prog + prog[H_OFF_RAM] + r * (ab + 1u + width);
    const uint64_t ctl_row = t <= ab ? P[((size_t)rp[t] << lg) + g] : 0ull;
    const uint64_t aw = bp_transpose64(ctl_row, t);     // bits [0, ab): address, bit ab: enable
    const uint64_t data_row = t < width ? P[((size_t)rp[ab + 1u + t] << lg) + g] : 0ull;
    const uint64_t wd = bp_transpose64(data_row, t);
    const uint32_t lane = g * 64u + t;
    const uint32_t addr = (uint32_t)(aw & ((1ull << ab) - 1ull));
    const bool we = (aw >> ab) & 1ull;
    const uint64_t wmask = width >= 64 ? ~0ull : ((1ull << width) - 1ull);
    uint64_t* mem = M + ((size_t)lane * nr + r) * depth;
    const uint64_t old = addr < depth ? mem[addr] : 0ull;
    if (we && addr < depth) mem[addr] = wd & wmask;
    const uint64_t col = bp_transpose64(old, t);
    if (t < width) P[((size_t)(n_pi + n_ff + r * width + t) << lg) + g] = col;
}

__global__ void __launch_bounds__(64) gn_bp_hash(const uint32_t* __restrict__ prog, const uint64_t* __restrict__ P,
                                                 uint32_t lg, uint32_t nlanes, uint64_t* __restrict__ hash,
                                                 const uint32_t* __restrict__ budget,
                                                 const uint32_t* __restrict__ ctl) {
    const uint32_t g = blockIdx.x, t = threadIdx.x;
    const uint32_t n_po = prog[H_PO];
    const uint32_t* po = prog + prog[H_OFF_PO];
    const uint64_t row = t < n_po ? P[((size_t)po[t] << lg) + g] : 0ull;
    const uint64_t v = bp_transpose64(row, t);
    const uint32_t lane = g * 64u + t;
    if (lane >= nlanes || ctl[0] >= budget[lane]) return;
    const int nb = n_po <= 8 ? 1 : n_po <= 16 ? 2 : n_po <= 32 ? 4 : 8;
    hash[lane] = rf_fnv_bytes(hash[lane], v, nb);
}

__global__ void gn_tick(uint32_t* ctl) {
    if (threadIdx.x == 0) ctl[0] += 1u;
}

__global__ void __launch_bounds__(256) gn_reset(uint64_t* hash, uint32_t n, uint32_t* ctl, uint64_t seed) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < n) hash[k] = RF_FNV_BASIS;
    if (k == 0) {
        ctl[0] = 0u;
        ctl[1] = 0u;
        ctl[2] = (uint32_t)seed;
        ctl[3] = (uint32_t)(seed >> 32);
    }
}

}  // namespace

struct rf_gatenet {
    std::vector<uint32_t> host;     // the program
    uint32_t nlanes = 0;
    uint32_t lg = 0;                // log2 of the bit-plane groups
    uint64_t npad = 0;              // per-lane form: lane stride
    hipStream_t stream = nullptr;
    uint32_t* prog = nullptr;
    uint64_t* P = nullptr;          // bit planes
    uint8_t* X = nullptr;           // per-lane bytes
    uint64_t* M = nullptr;          // RAM rows
    uint64_t* hash = nullptr;
    uint32_t* budget = nullptr;
    uint32_t* ctl = nullptr;
    uint32_t* sid = nullptr;        // bit-plane form: stimulus of each lane
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint32_t nodes[2] = {0, 0};
    // bit-plane cycle graphs over the first 2^e groups, e = 0 .. lg
    hipGraph_t bgraph[32] = {};
    hipGraphExec_t bexec[32] = {};
    uint32_t bnodes[32] = {};
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

const uint32_t* H(const rf_gatenet* g) { return g->host.data(); }

// Add a kernel node; ``args`` are the kernel's parameters (copied).
int add_kernel(hipGraph_t gr, const void* fn, dim3 grid, dim3 block, void** args, const hipGraphNode_t* deps,
               size_t ndeps, hipGraphNode_t* node) {
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(fn);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = args;
    p.extra = nullptr;
    GN_HIP(hipGraphAddKernelNode(node, gr, deps, ndeps, &p));
    return RF_OK;
}

// Levels of records [lev[l], lev[l + 1]) at ``rec``: one bit-plane gate
// kernel per level, chained after ``prev``.
int add_bp_levels(rf_gatenet* g, hipGraph_t gr, const uint32_t* lev, uint32_t nlev, const uint4* rec,
                  uint32_t le, hipGraphNode_t* prev, uint32_t* count) {
    for (uint32_t l = 0; l < nlev; ++l) {
        const uint32_t n = lev[l + 1] - lev[l];
        if (!n) continue;
        const uint4* r = rec + lev[l];
        uint64_t* P = g->P;
        uint32_t nn = n, lg = g->lg, e = le;
        void* args[] = {&r, &nn, &P, &lg, &e};
        const uint64_t threads = (uint64_t)n << le;
        hipGraphNode_t node;
        int rc = add_kernel(gr, (const void*)gn_bp_gates, dim3((uint32_t)((threads + 255) / 256)), dim3(256), args,
                            prev, 1, &node);
        if (rc) return rc;
        *prev = node;
        ++*count;
    }
    return RF_OK;
}

// The bit-plane cycle graph over the first E = 2^le groups.
int build_bitplane_graph(rf_gatenet* g, uint32_t le) {
    const uint32_t* h = H(g);
    hipGraph_t gr;
    GN_HIP(hipGraphCreate(&gr, 0));
    g->bgraph[le] = gr;
    const uint32_t G = 1u << le;
    uint32_t count = 0;
    hipGraphNode_t prev;
    {
        uint64_t* P = g->P;
        uint32_t lg = g->lg, npi = h[H_PI], nl = g->nlanes;
        uint32_t* ctl = g->ctl;
        const uint32_t* sid = g->sid;
        void* args[] = {&P, &lg, &npi, &ctl, &sid, &nl};
        int rc = add_kernel(gr, (const void*)gn_bp_stim, dim3(G), dim3(64), args, nullptr, 0, &prev);
        if (rc) return rc;
        ++count;
    }
    const uint4* recf = reinterpret_cast<const uint4*>(g->prog + h[H_OFF_REC_FULL]);
    const uint4* recc = reinterpret_cast<const uint4*>(g->prog + h[H_OFF_REC_CONE]);
    int rc = add_bp_levels(g, gr, h + h[H_OFF_LEV_FULL], h[H_LEV_FULL], recf, le, &prev, &count);
    if (rc) return rc;
    {
        const uint32_t ffb = (uint32_t)((((uint64_t)h[H_FF] << le) + 63) / 64);
        const uint32_t blocks = ffb + G * h[H_RAMS];
        const uint32_t* prog = g->prog;
        uint64_t* P = g->P;
        uint32_t lg = g->lg, e = le, fb = ffb;
        uint64_t* M = g->M;
        void* args[] = {&prog, &P, &lg, &e, &fb, &M};
        hipGraphNode_t node;
        if (blocks) {
            rc = add_kernel(gr, (const void*)gn_bp_edge, dim3(blocks), dim3(64), args, &prev, 1, &node);
            if (rc) return rc;
            prev = node;
            ++count;
        }
    }
    rc = add_bp_levels(g, gr, h + h[H_OFF_LEV_CONE], h[H_LEV_CONE], recc, le, &prev, &count);
    if (rc) return rc;
    {
        const uint32_t* prog = g->prog;
        const uint64_t* P = g->P;
        uint32_t lg = g->lg, nl = g->nlanes;
        uint64_t* hash = g->hash;
        const uint32_t* budget = g->budget;
        const uint32_t* ctl = g->ctl;
        void* args[] = {&prog, &P, &lg, &nl, &hash, &budget, &ctl};
        hipGraphNode_t node;
        rc = add_kernel(gr, (const void*)gn_bp_hash, dim3(G), dim3(64), args, &prev, 1, &node);
        if (rc) return rc;
        prev = node;
        ++count;
    }
    {
        uint32_t* ctl = g->ctl;
        void* args[] = {&ctl};
        hipGraphNode_t node;
        rc = add_kernel(gr, (const void*)gn_tick, dim3(1), dim3(64), args, &prev, 1, &node);
        if (rc) return rc;
        ++count;
    }
    GN_HIP(hipGraphInstantiate(&g->bexec[le], gr, nullptr, nullptr, 0));
    g->bnodes[le] = count;
    return RF_OK;
}

// One eval of the mtask DAG after ``dep``; *tail: a node after all sinks.
int add_mtask_dag(rf_gatenet* g, hipGraph_t gr, hipGraphNode_t dep, hipGraphNode_t* tail, uint32_t* count) {
    const uint32_t* h = H(g);
    const uint32_t nm = h[H_MTASKS];
    const uint32_t* mt = h + h[H_OFF_MT];
    const uint32_t* preds = h + h[H_OFF_PREDS];
    const uint4* rec = reinterpret_cast<const uint4*>(g->prog + h[H_OFF_REC_MT]);
    std::vector<hipGraphNode_t> nodes(nm);
    std::vector<char> has_succ(nm, 0);
    const dim3 grid((g->nlanes + 255) / 256);
    for (uint32_t k = 0; k < nm; ++k) {
        std::vector<hipGraphNode_t> deps;
        for (uint32_t j = 0; j < mt[4 * k + 2]; ++j) {
            const uint32_t p = preds[mt[4 * k + 3] + j];
            deps.push_back(nodes[p]);
            has_succ[p] = 1;
        }
        if (deps.empty()) deps.push_back(dep);
        const uint4* r = rec + mt[4 * k];
        uint32_t n = mt[4 * k + 1], nl = g->nlanes;
        uint8_t* X = g->X;
        uint64_t npad = g->npad;
        void* args[] = {&r, &n, &X, &npad, &nl};
        int rc = add_kernel(gr, (const void*)gn_lane_mtask, grid, dim3(256), args, deps.data(), deps.size(),
                            &nodes[k]);
        if (rc) return rc;
        ++*count;
    }
    std::vector<hipGraphNode_t> sinks;
    for (uint32_t k = 0; k < nm; ++k)
        if (!has_succ[k]) sinks.push_back(nodes[k]);
    if (sinks.empty()) sinks.push_back(dep);
    GN_HIP(hipGraphAddEmptyNode(tail, gr, sinks.data(), sinks.size()));
    return RF_OK;
}

int build_lane_graph(rf_gatenet* g) {
    const uint32_t* h = H(g);
    hipGraph_t gr;
    GN_HIP(hipGraphCreate(&gr, 0));
    g->graph[RF_GN_GRAPH] = gr;
    const dim3 grid((g->nlanes + 255) / 256);
    uint32_t count = 0;
    hipGraphNode_t stim, t1, edge, t2, hashn, tick;
    {
        uint8_t* X = g->X;
        uint64_t npad = g->npad;
        uint32_t nl = g->nlanes, npi = h[H_PI];
        const uint32_t* ctl = g->ctl;
        void* args[] = {&X, &npad, &nl, &npi, &ctl};
        int rc = add_kernel(gr, (const void*)gn_lane_stim, grid, dim3(256), args, nullptr, 0, &stim);
        if (rc) return rc;
        ++count;
    }
    int rc = add_mtask_dag(g, gr, stim, &t1, &count);
    if (rc) return rc;
    {
        const uint32_t* prog = g->prog;
        uint8_t* X = g->X;
        uint64_t npad = g->npad;
        uint32_t nl = g->nlanes;
        uint64_t* M = g->M;
        void* args[] = {&prog, &X, &npad, &nl, &M};
        rc = add_kernel(gr, (const void*)gn_lane_edge, grid, dim3(256), args, &t1, 1, &edge);
        if (rc) return rc;
        ++count;
    }
    rc = add_mtask_dag(g, gr, edge, &t2, &count);
    if (rc) return rc;
    {
        const uint32_t* prog = g->prog;
        const uint8_t* X = g->X;
        uint64_t npad = g->npad;
        uint32_t nl = g->nlanes;
        uint64_t* hash = g->hash;
        const uint32_t* budget = g->budget;
        const uint32_t* ctl = g->ctl;
        void* args[] = {&prog, &X, &npad, &nl, &hash, &budget, &ctl};
        rc = add_kernel(gr, (const void*)gn_lane_hash, grid, dim3(256), args, &t2, 1, &hashn);
        if (rc) return rc;
        ++count;
    }
    {
        uint32_t* ctl = g->ctl;
        void* args[] = {&ctl};
        rc = add_kernel(gr, (const void*)gn_tick, dim3(1), dim3(64), args, &hashn, 1, &tick);
        if (rc) return rc;
        ++count;
    }
    GN_HIP(hipGraphInstantiate(&g->exec[RF_GN_GRAPH], gr, nullptr, nullptr, 0));
    g->nodes[RF_GN_GRAPH] = count;
    return RF_OK;
}

// Structural checks of a program: every table inside the blob, every node
// id inside the node range, records 16-byte aligned.
int check_program(const uint32_t* h, size_t words) {
    if (words < H_WORDS || h[H_MAGIC] != kMagic || h[H_VERSION] != kVersion)
        return rf_set_error(RF_E_ARG, "not a gate-net program (magic or version)");
    if (h[H_TOTAL] != words) return rf_set_error(RF_E_ARG, "program size %zu, header says %u", words, h[H_TOTAL]);
    const uint64_t nodes = h[H_NODES];
    if ((uint64_t)h[H_PI] + h[H_FF] + h[H_RAMBITS] > nodes || nodes > kDstMask)
        return rf_set_error(RF_E_ARG, "node counts out of range");
    if (h[H_PI] > 128 || h[H_PO] > 64 || h[H_WIDTH] > 64 || h[H_ABITS] > 31 || h[H_ABITS] + 1 > 64)
        return rf_set_error(RF_E_ARG, "port widths out of range");
    if ((uint64_t)h[H_RAMS] * h[H_WIDTH] != h[H_RAMBITS]) return rf_set_error(RF_E_ARG, "RAM bits mismatch");
    auto inside = [&](uint64_t off, uint64_t n) { return off + n <= words; };
    if (!inside(h[H_OFF_LEV_FULL], h[H_LEV_FULL] + 1ull) || !inside(h[H_OFF_LEV_CONE], h[H_LEV_CONE] + 1ull) ||
        !inside(h[H_OFF_FF], h[H_FF]) || !inside(h[H_OFF_PO], h[H_PO]) ||
        !inside(h[H_OFF_RAM], (uint64_t)h[H_RAMS] * (h[H_ABITS] + 1 + h[H_WIDTH])) ||
        !inside(h[H_OFF_REC_FULL], 4ull * h[H_REC_FULL]) || !inside(h[H_OFF_REC_CONE], 4ull * h[H_REC_CONE]) ||
        !inside(h[H_OFF_MT], 4ull * h[H_MTASKS]) || !inside(h[H_OFF_REC_MT], 4ull * h[H_REC_MT]))
        return rf_set_error(RF_E_ARG, "a table lies outside the program");
    if ((h[H_OFF_REC_FULL] | h[H_OFF_REC_CONE] | h[H_OFF_REC_MT]) & 3u)
        return rf_set_error(RF_E_ARG, "records not 16-byte aligned");
    auto check_lev = [&](uint32_t off, uint32_t nlev, uint32_t nrec) {
        const uint32_t* lev = h + off;
        if (lev[0] != 0 || lev[nlev] != nrec) return false;
        for (uint32_t l = 0; l < nlev; ++l)
            if (lev[l] > lev[l + 1]) return false;
        return true;
    };
    if (!check_lev(h[H_OFF_LEV_FULL], h[H_LEV_FULL], h[H_REC_FULL]) ||
        !check_lev(h[H_OFF_LEV_CONE], h[H_LEV_CONE], h[H_REC_CONE]))
        return rf_set_error(RF_E_ARG, "level table out of order");
    auto check_rec = [&](uint32_t off, uint32_t n) {
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t* r = h + off + 4 * k;
            const uint32_t ty = r[0] >> 29;
            if (ty > 4 || (r[0] & kDstMask) >= nodes || r[1] >= nodes || r[2] >= nodes || r[3] >= nodes)
                return false;
            if ((r[0] & kDstMask) < h[H_PI] + h[H_FF] + h[H_RAMBITS]) return false;
        }
        return true;
    };
    if (!check_rec(h[H_OFF_REC_FULL], h[H_REC_FULL]) || !check_rec(h[H_OFF_REC_CONE], h[H_REC_CONE]) ||
        !check_rec(h[H_OFF_REC_MT], h[H_REC_MT]))
        return rf_set_error(RF_E_ARG, "a gate record names a node out of range");
    for (uint32_t i = 0; i < h[H_FF]; ++i)
        if (h[h[H_OFF_FF] + i] >= nodes) return rf_set_error(RF_E_ARG, "flip-flop input out of range");
    for (uint32_t i = 0; i < h[H_PO]; ++i)
        if (h[h[H_OFF_PO] + i] >= nodes) return rf_set_error(RF_E_ARG, "output out of range");
    for (uint64_t i = 0; i < (uint64_t)h[H_RAMS] * (h[H_ABITS] + 1 + h[H_WIDTH]); ++i)
        if (h[h[H_OFF_RAM] + i] >= nodes) return rf_set_error(RF_E_ARG, "RAM port out of range");
    const uint32_t* mt = h + h[H_OFF_MT];
    for (uint32_t k = 0; k < h[H_MTASKS]; ++k) {
        if ((uint64_t)mt[4 * k] + mt[4 * k + 1] > h[H_REC_MT] || !inside(h[H_OFF_PREDS] + (uint64_t)mt[4 * k + 3], mt[4 * k + 2]))
            return rf_set_error(RF_E_ARG, "mtask %u out of range", k);
        for (uint32_t j = 0; j < mt[4 * k + 2]; ++j)
            if (h[h[H_OFF_PREDS] + mt[4 * k + 3] + j] >= k)
                return rf_set_error(RF_E_ARG, "mtask %u: predecessors must come first", k);
    }
    return RF_OK;
}

}  // namespace

extern "C" {

int rf_gn_destroy(rf_gatenet* g) {
    if (!g) return RF_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (int k = 0; k < 2; ++k) {
        if (g->exec[k]) (void)hipGraphExecDestroy(g->exec[k]);
        if (g->graph[k]) (void)hipGraphDestroy(g->graph[k]);
    }
    for (int k = 0; k < 32; ++k) {
        if (g->bexec[k]) (void)hipGraphExecDestroy(g->bexec[k]);
        if (g->bgraph[k]) (void)hipGraphDestroy(g->bgraph[k]);
    }
    void* ps[] = {g->prog, g->P, g->X, g->M, g->hash, g->budget, g->ctl, g->sid};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (g->e0) (void)hipEventDestroy(g->e0);
    if (g->e1) (void)hipEventDestroy(g->e1);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return RF_OK;
}

int rf_gn_create(const uint32_t* program, size_t words, uint32_t n_lanes, rf_gatenet** out) {
    if (!program || !out) return rf_set_error(RF_E_ARG, "rf_gn_create: null argument");
    *out = nullptr;
    if (!n_lanes) return rf_set_error(RF_E_ARG, "rf_gn_create: n_lanes must be > 0");
    int rc = check_program(program, words);
    if (rc) return rc;
    rf_gatenet* g = new rf_gatenet();
    g->host.assign(program, program + words);
    g->nlanes = n_lanes;
    uint32_t groups = (n_lanes + 63) / 64, lg = 0;
    while ((1u << lg) < groups) ++lg;
    g->lg = lg;
    g->npad = ((uint64_t)n_lanes + 255) / 256 * 256;
    auto cleanup = [&](int code) {
        rf_gn_destroy(g);
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&g->e0);
    if (e == hipSuccess) e = hipEventCreate(&g->e1);
    if (e == hipSuccess) e = hipMalloc(&g->prog, words * 4);
    if (e == hipSuccess) e = hipMemcpy(g->prog, program, words * 4, hipMemcpyHostToDevice);
    const size_t lanes_alloc = (size_t)64 << lg;
    const uint32_t* h = program;
    const size_t mbytes = std::max<size_t>(8, lanes_alloc * h[H_RAMS] * h[H_DEPTH] * 8);
    if (e == hipSuccess) e = hipMalloc(&g->M, mbytes);
    if (e == hipSuccess) e = hipMalloc(&g->hash, (size_t)n_lanes * 8);
    if (e == hipSuccess) e = hipMalloc(&g->budget, (size_t)n_lanes * 4);
    if (e == hipSuccess) e = hipMalloc(&g->sid, (size_t)n_lanes * 4);
    if (e == hipSuccess) e = hipMalloc(&g->ctl, 64);
    if (e != hipSuccess) return cleanup(rf_set_error(RF_E_HIP, "rf_gn_create: %s", hipGetErrorString(e)));
    *out = g;
    return RF_OK;
}

uint32_t rf_gn_mtasks(const rf_gatenet* g) { return g ? H(g)[H_MTASKS] : 0; }

uint32_t rf_gn_nodes_per_cycle(const rf_gatenet* g, uint32_t mode) {
    if (!g || mode > RF_GN_BITPLANE) return 0;
    return mode == RF_GN_BITPLANE ? g->bnodes[g->lg] : g->nodes[mode];
}

int rf_gn_run(rf_gatenet* g, uint32_t mode, uint64_t seed, const uint32_t* budgets, uint64_t* out_hashes,
              float* ms) {
    if (!g || !budgets || !out_hashes) return rf_set_error(RF_E_ARG, "rf_gn_run: null argument");
    if (mode > RF_GN_BITPLANE) return rf_set_error(RF_E_ARG, "rf_gn_run: unknown mode %u", mode);
    const uint32_t* h = H(g);
    if (mode == RF_GN_GRAPH && !h[H_MTASKS]) return rf_set_error(RF_E_STATE, "program has no mtask partition");
    const uint32_t ncyc = *std::max_element(budgets, budgets + g->nlanes);
    // state of this mode, allocated on first use
    if (mode == RF_GN_BITPLANE && !g->P) {
        const size_t bytes = ((size_t)h[H_NODES] << g->lg) * 8;
        GN_HIP(hipMalloc(&g->P, bytes));
    }
    if (mode == RF_GN_GRAPH && !g->X) {
        const size_t bytes = (size_t)h[H_NODES] * g->npad;
        GN_HIP(hipMalloc(&g->X, bytes));
    }
    if (mode == RF_GN_GRAPH && !g->exec[mode]) {
        int rc = build_lane_graph(g);
        if (rc) return rc;
    }
    // bit-plane form: stimuli dealt to lanes longest budget first, so the
    // groups with lanes still running at cycle c are a prefix of A(c)
    // groups and cycle c runs the graph over the first 2^ceil(log2 A(c))
    std::vector<uint32_t> lane_sid, lane_budget, le_of;
    if (mode == RF_GN_BITPLANE) {
        const uint32_t n = g->nlanes;
        lane_sid.resize(n);
        std::iota(lane_sid.begin(), lane_sid.end(), 0u);
        std::stable_sort(lane_sid.begin(), lane_sid.end(),
                         [&](uint32_t x, uint32_t y) { return budgets[x] > budgets[y]; });
        lane_budget.resize(n);
        for (uint32_t l = 0; l < n; ++l) lane_budget[l] = budgets[lane_sid[l]];
        const uint32_t ng = (n + 63) / 64;
        std::vector<uint32_t> gmax(ng, 0);
        for (uint32_t l = 0; l < n; ++l) gmax[l / 64] = std::max(gmax[l / 64], lane_budget[l]);
        le_of.resize(ncyc);
        uint32_t active = ng;
        for (uint32_t c = 0; c < ncyc; ++c) {
            while (active > 1 && gmax[active - 1] <= c) --active;
            uint32_t e = 0;
            while ((1u << e) < active) ++e;
            le_of[c] = e;
            if (!g->bexec[e]) {
                int rc = build_bitplane_graph(g, e);
                if (rc) return rc;
            }
        }
    }
    // the reset state: zero nodes and RAMs, hash basis, cycle 0
    if (mode == RF_GN_BITPLANE)
        GN_HIP(hipMemsetAsync(g->P, 0, ((size_t)h[H_NODES] << g->lg) * 8, g->stream));
    else
        GN_HIP(hipMemsetAsync(g->X, 0, (size_t)h[H_NODES] * g->npad, g->stream));
    const size_t lanes_alloc = (size_t)64 << g->lg;
    if (h[H_RAMS]) GN_HIP(hipMemsetAsync(g->M, 0, lanes_alloc * h[H_RAMS] * h[H_DEPTH] * 8, g->stream));
    if (mode == RF_GN_BITPLANE) {
        GN_HIP(hipMemcpyAsync(g->budget, lane_budget.data(), (size_t)g->nlanes * 4, hipMemcpyHostToDevice,
                              g->stream));
        GN_HIP(hipMemcpyAsync(g->sid, lane_sid.data(), (size_t)g->nlanes * 4, hipMemcpyHostToDevice, g->stream));
    } else {
        GN_HIP(hipMemcpyAsync(g->budget, budgets, (size_t)g->nlanes * 4, hipMemcpyHostToDevice, g->stream));
    }
    hipLaunchKernelGGL(gn_reset, dim3((g->nlanes + 255) / 256), dim3(256), 0, g->stream, g->hash, g->nlanes,
                       g->ctl, seed);
    GN_HIP(hipGetLastError());
    GN_HIP(hipEventRecord(g->e0, g->stream));
    for (uint32_t c = 0; c < ncyc; ++c)
        GN_HIP(hipGraphLaunch(mode == RF_GN_BITPLANE ? g->bexec[le_of[c]] : g->exec[mode], g->stream));
    GN_HIP(hipEventRecord(g->e1, g->stream));
    GN_HIP(hipEventSynchronize(g->e1));
    if (ms) GN_HIP(hipEventElapsedTime(ms, g->e0, g->e1));
    if (mode == RF_GN_BITPLANE) {
        std::vector<uint64_t> lh(g->nlanes);
        GN_HIP(hipMemcpyAsync(lh.data(), g->hash, (size_t)g->nlanes * 8, hipMemcpyDeviceToHost, g->stream));
        GN_HIP(hipStreamSynchronize(g->stream));
        for (uint32_t l = 0; l < g->nlanes; ++l) out_hashes[lane_sid[l]] = lh[l];
    } else {
        GN_HIP(hipMemcpyAsync(out_hashes, g->hash, (size_t)g->nlanes * 8, hipMemcpyDeviceToHost, g->stream));
        GN_HIP(hipStreamSynchronize(g->stream));
    }
    return RF_OK;
}

}  // extern "C"
