// rf_runtime.hip -- HIP runtime for generated design libraries (librf_verilated.so).
//
// Replaces the reference's generated RF::RTLflow runtime (src/V3EmitC.cpp:5246-5500):
//  * hipMalloc'd lane-minor SoA state instead of cudaMallocManaged buffers
//    (:5351-5366), zero-initialised (the --x-initial 0 reset);
//  * the mtask DAG is built as a hipGraph with explicit kernel nodes and
//    dependency edges (the reference builds a Taskflow cudaFlow, :5389-5424)
//    and instantiated once;
//  * the per-cycle protocol (inputs, clk=0 eval, clk=1 eval, output hash) is one
//    captured graph replayed per cycle, or one fused per-lane kernel, with no
//    host round trip per eval (the reference syncs and reads change[0] through
//    managed memory once per settle iteration, :5481-5491);
//  * the batched form launches the design's persistent kernel over a device
//    work queue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <deque>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "rf_bitplane.h"
#include "rf_verilated.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define RF_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(RF_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_));         \
    } while (0)

const size_t kClassBytes[4] = {1, 2, 4, 8};

}  // namespace

// The error channel of every entry point in the library (rf_last_error).
int rf_set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

struct rf_model {
    const rf_design* d = nullptr;
    rf_config cfg{};
    uint32_t nlanes = 0;
    uint64_t npad = 0;
    uint32_t grid = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int ncu = 0;                       // compute units of the device
    uint8_t* queue_scratch = nullptr;  // this device's rf_queue_scratch (rf_submit_queue_static)
    int queue_scratch_err = 0;         // hipGetSymbolAddress's error, if it failed
    void* cls[4] = {nullptr, nullptr, nullptr, nullptr};
    uint8_t* mem = nullptr;
    uint32_t* change = nullptr;
    uint8_t* done = nullptr;
    uint64_t* hash = nullptr;
    uint32_t* cyc = nullptr;
    uint32_t* stim = nullptr;
    uint32_t* flags = nullptr;
    uint32_t* budget = nullptr;
    uint32_t* order = nullptr;
    uint64_t* out_hash = nullptr;
    uint32_t qcap = 0;
    uint8_t* images = nullptr;     // device table of per-memory image descriptors
    std::vector<void*> image_bufs;
    bool did_init = false;
    bool graphs_stale = false;     // the captured graphs hold an old seed
    hipGraphExec_t eval_exec = nullptr;
    hipGraph_t eval_graph = nullptr;
    hipGraphExec_t cycle_exec = nullptr;
    hipGraph_t cycle_graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    rf_stats stats{};
    rf_args base{};
    uint32_t* bp_prog = nullptr;   // device copy of the bit-plane program
    unsigned long long* bp_progress = nullptr;  // progress counter of a bit-plane launch
    int bp_per_cu[2] = {0, 0};     // bit-plane workgroups per CU (plain / stream kernel), 0 = not asked yet
    rf_bp_params bp{};
    // launches of rf_run_cycles not yet waited for: their HIP events are
    // folded into the statistics (and device flags checked) by the next call
    // that waits for the device
    struct pending {
        hipEvent_t e0, e1;
        uint32_t kind, count;
    };
    std::vector<pending> pend;
    std::vector<hipEvent_t> evfree;
    // the open resumable queue (rf_queue_open): device budget / order /
    // out_hash hold the whole queue, positions [0, qpos) have run
    struct queue_run {
        bool open = false;
        uint64_t seed = 0;
        uint32_t nstim = 0, pos = 0;
        std::vector<uint32_t> budgets, ord;
    } q;
};

namespace {

rf_args args_for(const rf_model* m, uint32_t phase) {
    rf_args a = m->base;
    a.phase = phase;
    return a;
}

// A stream or capture is set: the fused and persistent launches take the
// kernels compiled with the stream code.
bool streams_set(const rf_model* m) { return m->base.in.data || m->base.out.data; }

int launch(rf_model* m, const void* fn, const rf_args& a) {
    rf_args copy = a;
    void* params[] = {&copy};
    RF_HIP(hipLaunchKernel(fn, dim3(m->grid), dim3(m->d->block), params, 0, m->stream));
    return RF_OK;
}

// Add one kernel node over all lanes after ``deps``.
int add_node(rf_model* m, hipGraph_t g, const void* fn, const rf_args& a, const hipGraphNode_t* deps,
             size_t ndeps, std::deque<rf_args>& argstore, hipGraphNode_t* node) {
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(fn);
    p.gridDim = dim3(m->grid);
    p.blockDim = dim3(m->d->block);
    argstore.push_back(a);
    void* kp[1] = {&argstore.back()};
    p.kernelParams = kp;
    RF_HIP(hipGraphAddKernelNode(node, g, deps, ndeps, &p));
    return RF_OK;
}

// Append one eval (mtask DAG, settle_unroll times, each followed by the
// eval-end node) to a graph. deps: nodes the eval depends on. On return
// *tail is the last eval-end node. The final eval-end uses ``last_phase``.
// With ``settle_tail`` (designs with change-detected variables) the passes
// all end in RF_PHASE_EVAL, then the design's settle-tail kernel runs the
// lanes still changing to convergence or the converge limit, and one more
// eval-end node does the phase work.
int add_eval(rf_model* m, hipGraph_t g, hipGraphNode_t dep, uint32_t unroll, uint32_t last_phase,
             bool settle_tail, std::deque<rf_args>& argstore, hipGraphNode_t* tail) {
    const rf_design* d = m->d;
    hipGraphNode_t prev = dep;
    for (uint32_t it = 0; it < unroll; ++it) {
        std::vector<hipGraphNode_t> nodes(d->nmtasks);
        std::vector<char> has_succ(d->nmtasks, 0);
        for (uint32_t k = 0; k < d->nmtasks; ++k)
            for (uint32_t j = 0; j < d->mtasks[k].npreds; ++j) has_succ[d->mtasks[k].preds[j]] = 1;
        for (uint32_t k = 0; k < d->nmtasks; ++k) {
            const rf_mtask& mt = d->mtasks[k];
            std::vector<hipGraphNode_t> deps;
            for (uint32_t j = 0; j < mt.npreds; ++j) deps.push_back(nodes[mt.preds[j]]);
            if (deps.empty() && prev) deps.push_back(prev);
            hipKernelNodeParams p{};
            p.func = const_cast<void*>(mt.kernel);
            p.gridDim = dim3(m->grid);
            p.blockDim = dim3(d->block);
            p.sharedMemBytes = 0;
            argstore.push_back(args_for(m, RF_PHASE_EVAL));
            void** kp = new void*[1];
            kp[0] = &argstore.back();
            p.kernelParams = kp;
            p.extra = nullptr;
            RF_HIP(hipGraphAddKernelNode(&nodes[k], g, deps.data(), deps.size(), &p));
            delete[] kp;
        }
        std::vector<hipGraphNode_t> sinks;
        for (uint32_t k = 0; k < d->nmtasks; ++k)
            if (!has_succ[k]) sinks.push_back(nodes[k]);
        if (sinks.empty() && prev) sinks.push_back(prev);
        hipKernelNodeParams p{};
        p.func = const_cast<void*>(d->k_eval_end);
        p.gridDim = dim3(m->grid);
        p.blockDim = dim3(d->block);
        argstore.push_back(args_for(m, it + 1 == unroll && !settle_tail ? last_phase : RF_PHASE_EVAL));
        void* kp[1] = {&argstore.back()};
        p.kernelParams = kp;
        hipGraphNode_t end;
        RF_HIP(hipGraphAddKernelNode(&end, g, sinks.data(), sinks.size(), &p));
        prev = end;
    }
    if (settle_tail) {
        rf_args a = args_for(m, RF_PHASE_EVAL);
        a.ncycles = unroll;
        hipGraphNode_t st, end;
        int rc = add_node(m, g, d->k_settle, a, &prev, 1, argstore, &st);
        if (rc) return rc;
        rc = add_node(m, g, d->k_eval_end, args_for(m, last_phase), &st, 1, argstore, &end);
        if (rc) return rc;
        prev = end;
    }
    *tail = prev;
    return RF_OK;
}

int build_graphs(rf_model* m) {
    const rf_design* d = m->d;
    // argument blocks must outlive node creation only (kernel params are copied)
    std::deque<rf_args> argstore;
    // eval graph: one eval (used by rf_eval, one settle iteration per replay)
    RF_HIP(hipGraphCreate(&m->eval_graph, 0));
    hipGraphNode_t tail;
    int rc = add_eval(m, m->eval_graph, nullptr, 1, RF_PHASE_EVAL, false, argstore, &tail);
    if (rc) return rc;
    RF_HIP(hipGraphInstantiate(&m->eval_exec, m->eval_graph, nullptr, nullptr, 0));
    // cycle graph: stim -> eval(clk=0) -> rise -> eval(clk=1) -> hash
    RF_HIP(hipGraphCreate(&m->cycle_graph, 0));
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(d->k_stim);
    p.gridDim = dim3(m->grid);
    p.blockDim = dim3(d->block);
    argstore.push_back(args_for(m, RF_PHASE_EVAL));
    void* kp[1] = {&argstore.back()};
    p.kernelParams = kp;
    hipGraphNode_t stim;
    RF_HIP(hipGraphAddKernelNode(&stim, m->cycle_graph, nullptr, 0, &p));
    const uint32_t unroll = d->nchange ? std::max<uint32_t>(1, m->cfg.settle_unroll) : 1;
    const bool settle_tail = d->nchange && d->k_settle;
    hipGraphNode_t t1, t2;
    rc = add_eval(m, m->cycle_graph, stim, unroll, RF_PHASE_RISE, settle_tail, argstore, &t1);
    if (rc) return rc;
    rc = add_eval(m, m->cycle_graph, t1, unroll, RF_PHASE_HASH, settle_tail, argstore, &t2);
    if (rc) return rc;
    RF_HIP(hipGraphInstantiate(&m->cycle_exec, m->cycle_graph, nullptr, nullptr, 0));
    return RF_OK;
}

// Check the design's bit-plane program against the layout and copy it to
// the device.
int load_bitplane(rf_model* m) {
    const rf_design* d = m->d;
    const uint32_t* b = d->bitplane;
    const char* why = nullptr;
    if (rf_bp_parse(b, d->bitplane_words, &m->bp, &why)) return fail(RF_E_DESIGN, "%s", why);
    const rf_bp_params& p = m->bp;
    for (uint32_t k = 0; k < p.n_state; ++k) {
        const uint32_t* sl = b + p.off_state + 6 * k;
        const uint64_t loc = (uint64_t)sl[1] | ((uint64_t)sl[2] << 32);
        const uint64_t n = sl[4] > 64 ? sl[3] : 1;
        if (loc + n > d->slots[sl[0]]) return fail(RF_E_DESIGN, "bit-plane state slot outside the layout");
    }
    uint64_t q = p.off_f;
    for (uint32_t c = 0; c < p.n_f; ++c) {
        const uint32_t nl = b[q], ng = b[q + 1], nwb = b[q + 2];
        uint64_t e = q + RF_BP_FCHUNK_HDR + nl + 1;
        e += e & 1;
        e += 2ull * ng;
        for (uint32_t k = 0; k < nwb; ++k) {
            const uint64_t loc = (uint64_t)b[e + 1] | ((uint64_t)b[e + 2] << 32);
            const uint64_t n = b[e + 4] > 64 ? b[e + 3] : 1;
            if (loc + n > d->slots[b[e]])
                return fail(RF_E_DESIGN, "bit-plane write-back slot outside the layout");
            e += 5 + b[e + 4];
        }
        q += b[q + 4];
    }
    // memories: the blob's table must describe the design's own layout
    for (uint32_t k = 0; k < p.n_mem; ++k) {
        const uint32_t* me = b + p.off_mem + RF_BP_MEM_WORDS * k;
        if (me[0] >= d->nmems) return fail(RF_E_DESIGN, "bit-plane memory outside the design");
        const rf_mem& md = d->mems[me[0]];
        const uint64_t loc = (uint64_t)me[3] | ((uint64_t)me[4] << 32);
        if (md.cls != me[1] || md.layout != me[2] || md.loc != loc || md.depth != me[6] ||
            md.width != me[7] || md.words != 1 || (me[2] && md.row_bytes != me[5]))
            return fail(RF_E_DESIGN, "bit-plane memory %s does not match the design layout", md.name);
        if (me[2] ? loc + md.row_bytes > d->mem_row_bytes : loc + md.depth > d->slots[md.cls])
            return fail(RF_E_DESIGN, "bit-plane memory %s outside the layout", md.name);
    }
    const size_t bytes = d->bitplane_words * 4;
    RF_HIP(hipMalloc(&m->bp_prog, bytes));
    RF_HIP(hipMemcpy(m->bp_prog, b, bytes, hipMemcpyHostToDevice));
    m->bp.prog = m->bp_prog;
    RF_HIP(hipMalloc(&m->bp_progress, sizeof(unsigned long long)));
    m->bp.progress = m->bp_progress;
    return RF_OK;
}

enum { KIND_FUSED = 0, KIND_GRAPH = 1, KIND_BITPLANE = 2 };
constexpr size_t kMaxPending = 64;

int ev_take(rf_model* m, hipEvent_t* e) {
    if (!m->evfree.empty()) {
        *e = m->evfree.back();
        m->evfree.pop_back();
        return RF_OK;
    }
    RF_HIP(hipEventCreate(e));
    return RF_OK;
}

// Open a timed region on the model's stream (*e0 recorded).
int timed_begin(rf_model* m, hipEvent_t* e0) {
    int rc = ev_take(m, e0);
    if (rc) return rc;
    RF_HIP(hipEventRecord(*e0, m->stream));
    return RF_OK;
}

// Close it: ``count`` launches of ``kind`` are timed when next waited for.
int timed_end(rf_model* m, hipEvent_t e0, uint32_t kind, uint32_t count) {
    hipEvent_t e1;
    int rc = ev_take(m, &e1);
    if (rc) return rc;
    RF_HIP(hipEventRecord(e1, m->stream));
    m->pend.push_back({e0, e1, kind, count});
    return RF_OK;
}

int check_flags(rf_model* m);

// Queue buffers of rf_submit_queue_static: a file-scope device array, one
// copy per GPU for the life of the process. Layout: out_hash[max] (u64),
// budget[max] (u32), order[max] (u32). In host code ``rf_queue_scratch``
// names a host-side stand-in, never a device address: only the address that
// hipGetSymbolAddress returns for the current device may reach a kernel or
// a copy.
static __device__ __align__(256) uint8_t rf_queue_scratch[(size_t)RF_QUEUE_STATIC_MAX * 16];

constexpr size_t kScratchHashOff = 0;
constexpr size_t kScratchBudgetOff = (size_t)RF_QUEUE_STATIC_MAX * 8;
constexpr size_t kScratchOrderOff = (size_t)RF_QUEUE_STATIC_MAX * 12;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// This device's address of rf_queue_scratch, looked up once per device with
// that device current (called by rf_create on the model's device).
hipError_t queue_scratch_for(int dev, uint8_t** out) {
    *out = nullptr;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[dev]) {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(rf_queue_scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? (uint8_t*)p : nullptr;
        g_scratch_done[dev] = true;
    }
    *out = g_scratch_ptr[dev];
    return g_scratch_err[dev];
}

// Host staging of the static queue's order (no heap): guarded by its mutex.
std::mutex g_static_order_mu;
uint32_t g_static_order[RF_QUEUE_STATIC_MAX];
uint64_t g_static_seen[RF_QUEUE_STATIC_MAX / 64];

// rf_begin's per-lane protocol state, set on the device so that a begin
// queues behind earlier launches without a host wait.
__global__ void rf_rt_begin(uint64_t* hash, uint32_t* stim, uint32_t* cyc, uint64_t npad,
                            uint32_t stim0) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npad) return;
    hash[k] = 0xcbf29ce484222325ull;  // FNV-1a-64 offset basis
    stim[k] = stim0 + (uint32_t)k;
    cyc[k] = 0u;
}

void drop_graphs(rf_model* m) {
    if (m->cycle_exec) (void)hipGraphExecDestroy(m->cycle_exec);
    if (m->cycle_graph) (void)hipGraphDestroy(m->cycle_graph);
    if (m->eval_exec) (void)hipGraphExecDestroy(m->eval_exec);
    if (m->eval_graph) (void)hipGraphDestroy(m->eval_graph);
    m->cycle_exec = nullptr;
    m->cycle_graph = nullptr;
    m->eval_exec = nullptr;
    m->eval_graph = nullptr;
}

// Rebuild the graphs when their kernel arguments are out of date.
int fresh_graphs(rf_model* m) {
    if (!m->graphs_stale && m->cycle_exec) return RF_OK;
    drop_graphs(m);
    int rc = build_graphs(m);
    if (rc) return rc;
    m->graphs_stale = false;
    return RF_OK;
}

// Wait for every queued launch, fold the HIP-event times into the
// statistics and (``check``) report the run's device flags.
int settle(rf_model* m, bool check = true) {
    if (m->pend.empty()) return RF_OK;
    hipError_t e = hipEventSynchronize(m->pend.back().e1);
    for (const auto& q : m->pend) {
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, q.e0, q.e1);
        if (q.kind == KIND_FUSED) {
            m->stats.fused_launches += q.count;
            m->stats.fused_ms += ms;
        } else if (q.kind == KIND_GRAPH) {
            m->stats.graph_launches += q.count;
            m->stats.graph_ms += ms;
        } else {
            m->stats.bitplane_launches += q.count;
            m->stats.bitplane_ms += ms;
        }
        m->evfree.push_back(q.e0);
        m->evfree.push_back(q.e1);
    }
    m->pend.clear();
    if (e != hipSuccess) return fail(RF_E_HIP, "queued launches: %s", hipGetErrorString(e));
    return check ? check_flags(m) : RF_OK;
}

int do_init(rf_model* m) {
    if (m->did_init) return RF_OK;
    int rc = launch(m, m->d->k_init, args_for(m, RF_PHASE_EVAL));
    if (rc) return rc;
    m->did_init = true;
    return RF_OK;
}

int check_flags(rf_model* m) {
    uint32_t f[RF_FLAG_WORDS];
    RF_HIP(hipMemcpyAsync(f, m->flags, sizeof f, hipMemcpyDeviceToHost, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    if (f[RF_FLAG_UNCONVERGED])
        return fail(RF_E_CONVERGE, "Verilated model didn't converge (settle limit %u reached)",
                    m->cfg.converge_limit);
    if (f[RF_FLAG_STREAM]) {
        const rf_stream& in = m->base.in;
        return fail(RF_E_ARG, "a running lane read a stimulus-cycle outside the input stream "
                    "(stimuli [%llu, %llu), cycles [%u, %llu))", (unsigned long long)in.stim0,
                    (unsigned long long)(in.stim0 + in.nstim), in.cycle0,
                    (unsigned long long)in.cycle0 + in.ncycles);
    }
    return RF_OK;
}

// 64-bit words per stimulus-cycle of the input stream and the output capture.
void stream_words(const rf_design* d, uint32_t* nin, uint32_t* nout) {
    uint32_t a = 0, b = 0;
    for (uint32_t p = 0; p < d->nports; ++p) {
        const rf_port& q = d->ports[p];
        const uint32_t k = (q.width + 63) / 64;
        if (q.dir == RF_PORT_IN) a += k;
        else if (q.dir == RF_PORT_OUT && q.hashed) b += k;
    }
    *nin = a;
    *nout = b;
}

int set_stream(rf_model* m, rf_stream* dst, uint64_t* words, uint64_t stim0, uint32_t nstim, uint32_t cycle0,
               uint32_t ncycles, uint32_t nwords, const char* what) {
    if (!words) {
        *dst = rf_stream{};
        m->graphs_stale = true;
        return RF_OK;
    }
    if (!nstim || !ncycles) return fail(RF_E_ARG, "%s: empty window (%u stimuli, %u cycles)", what, nstim, ncycles);
    if (((uintptr_t)words & 7u) != 0) return fail(RF_E_ARG, "%s: buffer not 8-byte aligned", what);
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, words);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(RF_E_ARG, "%s: %p is not GPU memory", what, (void*)words);
    }
    const bool dev = at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged ||
                     at.type == hipMemoryTypeUnified;
    if (!dev) return fail(RF_E_ARG, "%s: %p is not device or managed memory", what, (void*)words);
    if (at.type == hipMemoryTypeDevice && at.device != m->device)
        return fail(RF_E_ARG, "%s: buffer on GPU %d, model on GPU %d", what, at.device, m->device);
    // drain queued launches that may still read the old buffer's settings
    // (they captured their arguments already; this only keeps the order of
    // errors plain)
    const int src = settle(m);
    if (src) return src;
    *dst = rf_stream{words, stim0, nstim, cycle0, ncycles, nwords};
    m->graphs_stale = true;
    return RF_OK;
}

// The batched forms hand out stimulus ids [0, nstim) for cycles [0, budget):
// the input stream must hold all of them.
int queue_stream_check(rf_model* m, const uint32_t* budgets, uint32_t nstim) {
    const rf_stream& in = m->base.in;
    if (!in.data) return RF_OK;
    uint32_t mx = 0;
    for (uint32_t k = 0; k < nstim; ++k) mx = std::max(mx, budgets[k]);
    if (in.stim0 != 0 || in.nstim < nstim || (mx && (in.cycle0 != 0 || in.ncycles < mx)))
        return fail(RF_E_ARG, "the input stream holds stimuli [%llu, %llu) and cycles [%u, %llu); "
                    "the queue needs stimuli [0, %u) and cycles [0, %u)", (unsigned long long)in.stim0,
                    (unsigned long long)(in.stim0 + in.nstim), in.cycle0,
                    (unsigned long long)in.cycle0 + in.ncycles, nstim, mx);
    return RF_OK;
}

}  // namespace

extern "C" {

uint32_t rf_abi_version(void) { return RF_DESIGN_ABI; }

const char* rf_last_error(void) { return g_err.c_str(); }

uint32_t rf_lanes(const rf_model* m) { return m ? m->nlanes : 0; }
uint64_t rf_lane_stride(const rf_model* m) { return m ? m->npad : 0; }

int rf_create(const rf_design* d, uint32_t n_lanes, const rf_config* cfg, rf_model** out) {
    if (!d || !out) return fail(RF_E_ARG, "rf_create: null design or output pointer");
    *out = nullptr;
    if (d->abi != RF_DESIGN_ABI)
        return fail(RF_E_DESIGN, "design ABI %u, runtime expects %u", d->abi, RF_DESIGN_ABI);
    if (n_lanes == 0) return fail(RF_E_ARG, "rf_create: n_lanes must be > 0");
    if (cfg && cfg->bitplane_threads &&
        !(cfg->bitplane_threads == 128 || cfg->bitplane_threads == 256 || cfg->bitplane_threads == 512 ||
          cfg->bitplane_threads == 1024))
        return fail(RF_E_ARG, "bitplane_threads must be 128, 256, 512 or 1024 (got %u)", cfg->bitplane_threads);
    if (cfg && cfg->lane_pad % 64)
        return fail(RF_E_ARG, "lane_pad must be a multiple of 64 (got %u)", cfg->lane_pad);
    rf_model* m = new rf_model();
    m->d = d;
    m->cfg.settle_unroll = 2;
    m->cfg.converge_limit = 100;
    m->cfg.fused_chunk = 512;
    m->cfg.bitplane_threads = RF_BP_DEFAULT_THREADS;
    if (cfg) {
        if (cfg->settle_unroll) m->cfg.settle_unroll = cfg->settle_unroll;
        if (cfg->converge_limit) m->cfg.converge_limit = cfg->converge_limit;
        if (cfg->fused_chunk) m->cfg.fused_chunk = cfg->fused_chunk;
        if (cfg->bitplane_threads) m->cfg.bitplane_threads = cfg->bitplane_threads;
        m->cfg.lane_pad = cfg->lane_pad;
        m->cfg.bitplane_slots = cfg->bitplane_slots;
        m->cfg.fused_multiwave = cfg->fused_multiwave;
    }
    m->nlanes = n_lanes;
    const uint64_t q = d->block > 64 ? d->block : 64;
    m->npad = ((uint64_t)n_lanes + q - 1) / q * q + m->cfg.lane_pad;
    m->grid = (uint32_t)((n_lanes + d->block - 1) / d->block);
    auto cleanup = [&](int rc) {
        rf_destroy(m);
        return rc;
    };
    hipError_t e = hipGetDevice(&m->device);
    if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipGetDevice: %s", hipGetErrorString(e)));
    e = hipDeviceGetAttribute(&m->ncu, hipDeviceAttributeMultiprocessorCount, m->device);
    if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipDeviceGetAttribute: %s", hipGetErrorString(e)));
    m->queue_scratch_err = (int)queue_scratch_for(m->device, &m->queue_scratch);
    // the bookkeeping of queued launches never grows past these, so waiting
    // for them (rf_submit_queue_static included) allocates nothing
    m->pend.reserve(kMaxPending + 1);
    m->evfree.reserve(2 * kMaxPending + 2);
    e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));
    for (int c = 0; c < 4; ++c) {
        const size_t bytes = std::max<uint64_t>(1, d->slots[c]) * m->npad * kClassBytes[c];
        e = hipMalloc(&m->cls[c], bytes);
        if (e != hipSuccess)
            return cleanup(fail(RF_E_HIP, "hipMalloc(%zu bytes) for class %d state: %s", bytes, c,
                                hipGetErrorString(e)));
        e = hipMemsetAsync(m->cls[c], 0, bytes, m->stream);
        if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipMemset: %s", hipGetErrorString(e)));
    }
    const size_t mbytes = std::max<uint64_t>(8, d->mem_row_bytes) * m->npad;
    struct {
        void** p;
        size_t bytes;
        int fill;
    } bufs[] = {
        {(void**)&m->mem, mbytes, 0},
        {(void**)&m->change, m->npad * 4, 1},
        {(void**)&m->done, m->npad, 0},
        {(void**)&m->hash, m->npad * 8, 0},
        {(void**)&m->cyc, m->npad * 4, 0},
        {(void**)&m->stim, m->npad * 4, 0},
        {(void**)&m->flags, RF_FLAG_WORDS * 4, 0},
    };
    for (auto& b : bufs) {
        e = hipMalloc(b.p, b.bytes);
        if (e != hipSuccess)
            return cleanup(fail(RF_E_HIP, "hipMalloc(%zu bytes): %s", b.bytes, hipGetErrorString(e)));
        e = hipMemsetAsync(*b.p, 0, b.bytes, m->stream);
        if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipMemset: %s", hipGetErrorString(e)));
    }
    // change[] = 1 for every lane (reference: cudaMemset(change, 1), src/V3EmitC.cpp:5363)
    {
        std::vector<uint32_t> ones(m->npad, 1u);
        e = hipMemcpyAsync(m->change, ones.data(), m->npad * 4, hipMemcpyHostToDevice, m->stream);
        if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipMemcpy: %s", hipGetErrorString(e)));
        e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "sync: %s", hipGetErrorString(e)));
    }
    e = hipMalloc(&m->images, (size_t)std::max<uint32_t>(1, d->nmems) * 16);
    if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipMalloc: %s", hipGetErrorString(e)));
    e = hipMemsetAsync(m->images, 0, (size_t)std::max<uint32_t>(1, d->nmems) * 16, m->stream);
    if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipMemset: %s", hipGetErrorString(e)));
    m->image_bufs.assign(d->nmems, nullptr);
    rf_args& a = m->base;
    a.c = (uint8_t*)m->cls[0];
    a.s = (uint16_t*)m->cls[1];
    a.i = (uint32_t*)m->cls[2];
    a.q = (uint64_t*)m->cls[3];
    a.mem = m->mem;
    a.change = m->change;
    a.done = m->done;
    a.hash = m->hash;
    a.cyc = m->cyc;
    a.stim = m->stim;
    a.flags = m->flags;
    a.images = nullptr;
    a.npad = m->npad;
    a.nlanes = n_lanes;
    a.converge_limit = m->cfg.converge_limit;
    a.ncycles = 0;
    e = hipEventCreate(&m->ev0);
    if (e == hipSuccess) e = hipEventCreate(&m->ev1);
    if (e != hipSuccess) return cleanup(fail(RF_E_HIP, "hipEventCreate: %s", hipGetErrorString(e)));
    int rc = build_graphs(m);
    if (rc) return cleanup(rc);
    if (d->bitplane && d->bitplane_words) {
        rc = load_bitplane(m);
        if (rc) return cleanup(rc);
    }
    *out = m;
    return RF_OK;
}

int rf_destroy(rf_model* m) {
    if (!m) return RF_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (const auto& q : m->pend) {
        (void)hipEventDestroy(q.e0);
        (void)hipEventDestroy(q.e1);
    }
    for (hipEvent_t e : m->evfree) (void)hipEventDestroy(e);
    drop_graphs(m);
    for (int c = 0; c < 4; ++c)
        if (m->cls[c]) (void)hipFree(m->cls[c]);
    void* ps[] = {m->mem, m->change, m->done, m->hash, m->cyc, m->stim, m->flags,
                  m->budget, m->order, m->out_hash, m->images, m->bp_prog, m->bp_progress};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (void* p : m->image_bufs)
        if (p) (void)hipFree(p);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return RF_OK;
}

static int port_check(rf_model* m, uint32_t port, uint32_t lane0, uint32_t count, const void* buf) {
    if (!m || !buf) return fail(RF_E_ARG, "null model or buffer");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (port >= m->d->nports) return fail(RF_E_ARG, "port %u out of range (%u ports)", port, m->d->nports);
    if ((uint64_t)lane0 + count > m->nlanes)
        return fail(RF_E_ARG, "lanes [%u, %u) exceed %u lanes", lane0, lane0 + count, m->nlanes);
    return RF_OK;
}

int rf_set_port(rf_model* m, uint32_t port, uint32_t lane0, uint32_t count, const void* src) {
    int rc = port_check(m, port, lane0, count, src);
    if (rc) return rc;
    if (!count) return RF_OK;
    const rf_port& p = m->d->ports[port];
    const size_t es = kClassBytes[p.cls];
    std::vector<uint8_t> stage((size_t)count * es);
    for (uint32_t w = 0; w < p.words; ++w) {
        for (uint32_t k = 0; k < count; ++k)
            memcpy(&stage[k * es], (const uint8_t*)src + ((size_t)k * p.words + w) * es, es);
        uint8_t* dst = (uint8_t*)m->cls[p.cls] + ((p.loc + w) * m->npad + lane0) * es;
        RF_HIP(hipMemcpyAsync(dst, stage.data(), (size_t)count * es, hipMemcpyHostToDevice, m->stream));
        RF_HIP(hipStreamSynchronize(m->stream));
    }
    return RF_OK;
}

int rf_get_port(rf_model* m, uint32_t port, uint32_t lane0, uint32_t count, void* dst) {
    int rc = port_check(m, port, lane0, count, dst);
    if (rc) return rc;
    if (!count) return RF_OK;
    const rf_port& p = m->d->ports[port];
    const size_t es = kClassBytes[p.cls];
    std::vector<uint8_t> stage((size_t)count * es);
    for (uint32_t w = 0; w < p.words; ++w) {
        const uint8_t* srcp = (const uint8_t*)m->cls[p.cls] + ((p.loc + w) * m->npad + lane0) * es;
        RF_HIP(hipMemcpyAsync(stage.data(), srcp, (size_t)count * es, hipMemcpyDeviceToHost, m->stream));
        RF_HIP(hipStreamSynchronize(m->stream));
        for (uint32_t k = 0; k < count; ++k)
            memcpy((uint8_t*)dst + ((size_t)k * p.words + w) * es, &stage[k * es], es);
    }
    return RF_OK;
}

static int mem_xfer(rf_model* m, uint32_t mi, uint32_t lane0, uint32_t count, uint32_t elem0,
                    uint32_t nelem, void* host, bool to_dev) {
    if (!m || !host) return fail(RF_E_ARG, "null model or buffer");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (mi >= m->d->nmems) return fail(RF_E_ARG, "memory %u out of range", mi);
    const rf_mem& md = m->d->mems[mi];
    if ((uint64_t)lane0 + count > m->nlanes) return fail(RF_E_ARG, "lane range out of bounds");
    if ((uint64_t)elem0 + nelem > md.depth) return fail(RF_E_ARG, "element range out of bounds");
    const size_t es = kClassBytes[md.cls];
    const size_t per_lane = (size_t)nelem * md.words * es;
    if (md.layout == RF_MEM_LANE_MAJOR) {
        for (uint32_t k = 0; k < count; ++k) {
            uint8_t* dev = m->mem + md.loc * m->npad + (size_t)(lane0 + k) * md.row_bytes +
                           (size_t)elem0 * md.words * es;
            uint8_t* h = (uint8_t*)host + k * per_lane;
            if (to_dev)
                RF_HIP(hipMemcpyAsync(dev, h, per_lane, hipMemcpyHostToDevice, m->stream));
            else
                RF_HIP(hipMemcpyAsync(h, dev, per_lane, hipMemcpyDeviceToHost, m->stream));
        }
        RF_HIP(hipStreamSynchronize(m->stream));
        return RF_OK;
    }
    std::vector<uint8_t> stage((size_t)count * es);
    for (uint32_t e = 0; e < nelem; ++e) {
        for (uint32_t w = 0; w < md.words; ++w) {
            const uint64_t slot = md.loc + (uint64_t)(elem0 + e) * md.words + w;
            uint8_t* dev = (uint8_t*)m->cls[md.cls] + (slot * m->npad + lane0) * es;
            if (to_dev) {
                for (uint32_t k = 0; k < count; ++k)
                    memcpy(&stage[k * es], (uint8_t*)host + k * per_lane + ((size_t)e * md.words + w) * es, es);
                RF_HIP(hipMemcpyAsync(dev, stage.data(), (size_t)count * es, hipMemcpyHostToDevice, m->stream));
                RF_HIP(hipStreamSynchronize(m->stream));
            } else {
                RF_HIP(hipMemcpyAsync(stage.data(), dev, (size_t)count * es, hipMemcpyDeviceToHost, m->stream));
                RF_HIP(hipStreamSynchronize(m->stream));
                for (uint32_t k = 0; k < count; ++k)
                    memcpy((uint8_t*)host + k * per_lane + ((size_t)e * md.words + w) * es, &stage[k * es], es);
            }
        }
    }
    return RF_OK;
}

int rf_mem_write(rf_model* m, uint32_t mem, uint32_t lane0, uint32_t count, uint32_t elem0,
                 uint32_t nelem, const void* src) {
    return mem_xfer(m, mem, lane0, count, elem0, nelem, const_cast<void*>(src), true);
}

int rf_mem_read(rf_model* m, uint32_t mem, uint32_t lane0, uint32_t count, uint32_t elem0,
                uint32_t nelem, void* dst) {
    return mem_xfer(m, mem, lane0, count, elem0, nelem, dst, false);
}

int rf_set_done(rf_model* m, uint32_t lane0, uint32_t count, const uint8_t* done) {
    if (!m || !done) return fail(RF_E_ARG, "null model or buffer");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if ((uint64_t)lane0 + count > m->nlanes) return fail(RF_E_ARG, "lane range out of bounds");
    RF_HIP(hipMemcpyAsync(m->done + lane0, done, count, hipMemcpyHostToDevice, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    return RF_OK;
}

int rf_eval(rf_model* m) {
    if (!m) return fail(RF_E_ARG, "null model");
    {
        const int src = settle(m);
        if (src) return src;
    }
    int rc = do_init(m);
    if (rc) return rc;
    rc = fresh_graphs(m);
    if (rc) return rc;
    const rf_design* d = m->d;
    if (!d->nchange) {
        RF_HIP(hipGraphLaunch(m->eval_exec, m->stream));
        RF_HIP(hipStreamSynchronize(m->stream));
        m->stats.graph_launches++;
        return RF_OK;
    }
    // settle loop driven from the host, one graph replay per iteration,
    // exactly the reference's condition-task loop (src/V3EmitC.cpp:5474-5491)
    for (uint32_t it = 0;; ++it) {
        RF_HIP(hipMemsetAsync(m->flags + RF_FLAG_CHANGE, 0, 4, m->stream));
        RF_HIP(hipGraphLaunch(m->eval_exec, m->stream));
        uint32_t ch = 0;
        RF_HIP(hipMemcpyAsync(&ch, m->flags + RF_FLAG_CHANGE, 4, hipMemcpyDeviceToHost, m->stream));
        RF_HIP(hipStreamSynchronize(m->stream));
        m->stats.graph_launches++;
        if (!ch) break;
        if (it + 1 >= m->cfg.converge_limit)
            return fail(RF_E_CONVERGE, "Verilated model didn't converge (settle limit %u reached)",
                        m->cfg.converge_limit);
    }
    // reference end_t: change[] = 1 for the next eval (src/V3EmitC.cpp:5477-5480)
    std::vector<uint32_t> ones(m->npad, 1u);
    RF_HIP(hipMemcpyAsync(m->change, ones.data(), m->npad * 4, hipMemcpyHostToDevice, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    return RF_OK;
}

int rf_begin(rf_model* m, uint64_t seed, uint32_t stim0) {
    if (!m) return fail(RF_E_ARG, "null model");
    {
        const int src = settle(m);
        if (src) return src;
    }
    const uint32_t tpb = 256;
    hipLaunchKernelGGL(rf_rt_begin, dim3((uint32_t)((m->npad + tpb - 1) / tpb)), dim3(tpb), 0, m->stream,
                       m->hash, m->stim, m->cyc, m->npad, stim0);
    RF_HIP(hipGetLastError());
    RF_HIP(hipMemsetAsync(m->flags, 0, RF_FLAG_WORDS * 4, m->stream));
    // the graphs captured the old seed: rebuilt before their next replay
    if (seed != m->base.seed) m->graphs_stale = true;
    m->base.seed = seed;
    return RF_OK;
}

// One launch of the bit-plane group rotation: a window of n consecutive
// groups (cyclically) from group base; the first split of them run c_lo
// cycles, the rest c_hi; f_lo / f_hi: those groups run for the last time
// and write back every stored variable (otherwise only their state).
struct bp_window {
    uint32_t base, n, split, c_lo, c_hi, f_lo, f_hi;
};

// Plan of the bit-plane group rotation for G groups of which S run at once,
// advancing every group n cycles. Each group runs in r of the launches
// (n / r cycles each, the first n % r of its runs one cycle more): the
// launches take consecutive windows of S positions of the sequence that
// lists groups 0..G-1 r times, so all but the last launch are full rounds
// and a group's runs come in order, each in a later launch. r is chosen by
// an estimate in cycle units of a full round: a launch costs its cycles
// plus a fixed 32 (state load, first and last programs, write-back); a
// single launch over several rounds costs n for the first round and 0.7 n
// for each further one (later rounds hold fewer workgroups per CU and run
// faster per cycle). forced: rotate whenever G > S. An empty plan means one
// launch over all groups.
static std::vector<bp_window> bp_rotation(uint32_t G, uint32_t S, uint32_t n, bool forced) {
    std::vector<bp_window> plan;
    if (!S || G <= S || !n) return plan;
    constexpr double kFixed = 32.0;
    const uint32_t rounds = (G + S - 1) / S;
    double best = forced ? 1e300 : n * (1.0 + 0.7 * (rounds - 1)) + kFixed;
    uint32_t br = 0;
    for (uint32_t r = 1; r <= 8 && r <= n; ++r) {
        const uint64_t launches = ((uint64_t)r * G + S - 1) / S;
        const double cost = (double)launches * ((n + r - 1) / r + kFixed);
        if (cost < best) best = cost, br = r;
    }
    if (!br) return plan;
    const uint64_t total = (uint64_t)br * G;
    auto cyc = [&](uint64_t j) { return n / br + (j < n % br ? 1u : 0u); };
    for (uint64_t q0 = 0; q0 < total; q0 += S) {
        bp_window w;
        w.n = (uint32_t)std::min<uint64_t>(S, total - q0);
        const uint64_t j0 = q0 / G;
        w.base = (uint32_t)(q0 % G);
        w.split = (uint32_t)std::min<uint64_t>(w.n, (j0 + 1) * G - q0);
        w.c_lo = cyc(j0);
        w.c_hi = j0 + 1 < br ? cyc(j0 + 1) : 0u;
        w.f_lo = j0 + 1 == br;
        w.f_hi = j0 + 2 == br;
        plan.push_back(w);
    }
    return plan;
}

int rf_run_cycles(rf_model* m, uint32_t n, uint32_t mode) {
    if (!m) return fail(RF_E_ARG, "null model");
    if (mode > RF_MODE_BITPLANE) return fail(RF_E_ARG, "unknown run mode %u", mode);
    if (mode == RF_MODE_BITPLANE && !m->bp_prog)
        return fail(RF_E_STATE, "design %s has no bit-plane program", m->d->name);
    int rc = m->pend.size() >= kMaxPending ? settle(m) : RF_OK;
    if (rc) return rc;
    rc = do_init(m);
    if (rc) return rc;
    hipEvent_t e0;
    if (mode == RF_MODE_GRAPH) {
        if (!n) return RF_OK;
        rc = fresh_graphs(m);
        if (rc) return rc;
        rc = timed_begin(m, &e0);
        if (rc) return rc;
        for (uint32_t c = 0; c < n; ++c) RF_HIP(hipGraphLaunch(m->cycle_exec, m->stream));
        return timed_end(m, e0, KIND_GRAPH, n);
    }
    uint32_t left = n;
    while (left) {
        const uint32_t chunk = std::min(left, m->cfg.fused_chunk);
        rc = timed_begin(m, &e0);
        if (rc) return rc;
        if (mode == RF_MODE_FUSED) {
            rf_args a = args_for(m, RF_PHASE_EVAL);
            a.ncycles = chunk;
            if (m->cfg.fused_multiwave && m->d->k_fused_mw && m->d->mw_waves >= 2 && m->d->mw_waves <= 16 &&
                !streams_set(m)) {
                // one workgroup of mw_waves wavefronts per 64 stimuli
                void* params[] = {&a};
                RF_HIP(hipLaunchKernel(m->d->k_fused_mw, dim3((m->nlanes + 63u) / 64u), dim3(64u * m->d->mw_waves),
                                       params, 0, m->stream));
            } else {
                rc = launch(m, streams_set(m) ? m->d->k_fused_io : m->d->k_fused, a);
                if (rc) return rc;
            }
        }
        uint32_t launches = 1;
        if (mode == RF_MODE_BITPLANE) {
            rf_bp_params p = m->bp;
            rf_bp_rounds(m->d->bitplane, m->cfg.bitplane_threads, &p);
            p.a = args_for(m, RF_PHASE_EVAL);
            p.a.ncycles = chunk;
            // groups beyond what the GPU holds at once rotate through
            // launches of one full round each
            const uint32_t G = (m->nlanes + 63u) / 64u;
            uint32_t S = m->cfg.bitplane_slots;
            if (!S) {
                int& per_cu = m->bp_per_cu[(p.a.in.data || p.a.out.data) ? 1 : 0];
                if (!per_cu) {
                    RF_HIP(rf_bp_occupancy(p, m->cfg.bitplane_threads, &per_cu));
                    per_cu = std::max(1, per_cu);
                }
                S = (uint32_t)per_cu * (uint32_t)std::max(1, m->ncu);
            }
            const std::vector<bp_window> plan = bp_rotation(G, S, chunk, m->cfg.bitplane_slots != 0);
            if (plan.empty()) {
                RF_HIP(hipMemsetAsync(m->bp_progress, 0, sizeof(unsigned long long), m->stream));
                RF_HIP(rf_bp_launch(p, m->cfg.bitplane_threads, m->stream));
            } else {
                p.grp_count = G;
                for (const bp_window& w : plan) {
                    p.grp_base = w.base;
                    p.grp_n = w.n;
                    p.grp_split = w.split;
                    p.a.ncycles = w.c_lo;
                    p.ncyc_hi = w.c_hi;
                    p.fwb_lo = w.f_lo;
                    p.fwb_hi = w.f_hi;
                    RF_HIP(hipMemsetAsync(m->bp_progress, 0, sizeof(unsigned long long), m->stream));
                    RF_HIP(rf_bp_launch(p, m->cfg.bitplane_threads, m->stream));
                }
                launches = (uint32_t)plan.size();
            }
        }
        rc = timed_end(m, e0, mode == RF_MODE_FUSED ? KIND_FUSED : KIND_BITPLANE, launches);
        if (rc) return rc;
        left -= chunk;
    }
    return RF_OK;
}

int rf_hashes(rf_model* m, uint32_t lane0, uint32_t count, uint64_t* out) {
    if (!m || !out) return fail(RF_E_ARG, "null model or buffer");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if ((uint64_t)lane0 + count > m->nlanes) return fail(RF_E_ARG, "lane range out of bounds");
    RF_HIP(hipMemcpyAsync(out, m->hash + lane0, (size_t)count * 8, hipMemcpyDeviceToHost, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    return RF_OK;
}

int rf_stream_words(const rf_design* d, uint32_t* in_words, uint32_t* out_words) {
    if (!d || !in_words || !out_words) return fail(RF_E_ARG, "null argument");
    stream_words(d, in_words, out_words);
    return RF_OK;
}

int rf_set_input_stream(rf_model* m, const uint64_t* words, uint64_t stim0, uint32_t nstim, uint32_t cycle0,
                        uint32_t ncycles) {
    if (!m) return fail(RF_E_ARG, "null model");
    uint32_t nin, nout;
    stream_words(m->d, &nin, &nout);
    return set_stream(m, &m->base.in, const_cast<uint64_t*>(words), stim0, nstim, cycle0, ncycles, nin,
                      "rf_set_input_stream");
}

int rf_set_output_capture(rf_model* m, uint64_t* words, uint64_t stim0, uint32_t nstim, uint32_t cycle0,
                          uint32_t ncycles) {
    if (!m) return fail(RF_E_ARG, "null model");
    uint32_t nin, nout;
    stream_words(m->d, &nin, &nout);
    return set_stream(m, &m->base.out, words, stim0, nstim, cycle0, ncycles, nout, "rf_set_output_capture");
}

// ---- resumable queue: the batched workload run in steps of queue
// positions; between steps no lane holds a stimulus, so the queue's state is
// its order, budgets, position and the hashes so far (rf_queue_save)

static int queue_setup(rf_model* m, const uint32_t* budgets, uint32_t nstim, const uint32_t* order,
                       std::vector<uint32_t>& ord);
static int bitplane_rounds(rf_model* m, uint64_t seed, const uint32_t* budgets, const uint32_t* ord,
                           uint32_t n, uint32_t nstim);

int rf_queue_open(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim, const uint32_t* order) {
    if (!m || (!budgets && nstim)) return fail(RF_E_ARG, "null argument");
    {
        const int src = settle(m);
        if (src) return src;
    }
    m->q = rf_model::queue_run{};
    int rc = queue_stream_check(m, budgets, nstim);
    if (rc) return rc;
    std::vector<uint32_t> ord;
    rc = queue_setup(m, budgets, nstim, order, ord);
    if (rc) return rc;
    if (nstim) {
        RF_HIP(hipMemcpyAsync(m->budget, budgets, (size_t)nstim * 4, hipMemcpyHostToDevice, m->stream));
        RF_HIP(hipMemcpyAsync(m->order, ord.data(), (size_t)nstim * 4, hipMemcpyHostToDevice, m->stream));
        RF_HIP(hipMemsetAsync(m->out_hash, 0, (size_t)nstim * 8, m->stream));
        RF_HIP(hipStreamSynchronize(m->stream));
    }
    m->q.open = true;
    m->q.seed = seed;
    m->q.nstim = nstim;
    m->q.budgets.assign(budgets, budgets + nstim);
    m->q.ord = std::move(ord);
    return RF_OK;
}

int rf_queue_step(rf_model* m, uint32_t max_stimuli, uint32_t mode, uint32_t* position) {
    if (!m) return fail(RF_E_ARG, "null model");
    if (!m->q.open) return fail(RF_E_STATE, "no open queue (rf_queue_open or rf_queue_load first)");
    if (mode != RF_MODE_FUSED && mode != RF_MODE_BITPLANE)
        return fail(RF_E_ARG, "queue steps run in RF_MODE_FUSED or RF_MODE_BITPLANE (got %u)", mode);
    {
        const int src = settle(m);
        if (src) return src;
    }
    auto& q = m->q;
    const uint32_t n = std::min(max_stimuli, q.nstim - q.pos);
    int rc = RF_OK;
    if (n && mode == RF_MODE_BITPLANE) {
        if (!m->bp_prog) return fail(RF_E_STATE, "design %s has no bit-plane program", m->d->name);
        if (m->base.images)
            return fail(RF_E_STATE, "memory images are loaded by the per-lane batched form only");
        RF_HIP(hipMemsetAsync(m->flags, 0, RF_FLAG_WORDS * 4, m->stream));
        rc = bitplane_rounds(m, q.seed, q.budgets.data(), q.ord.data() + q.pos, n, q.nstim);
    } else if (n) {
        uint32_t f[RF_FLAG_WORDS] = {0};
        f[RF_FLAG_QUEUE] = m->nlanes;
        RF_HIP(hipMemcpyAsync(m->flags, f, sizeof f, hipMemcpyHostToDevice, m->stream));
        rf_args a = args_for(m, RF_PHASE_EVAL);
        a.seed = q.seed;
        a.budget = m->budget;
        a.order = m->order + q.pos;   // this step's queue positions
        a.out_hash = m->out_hash;
        a.nstim = n;
        RF_HIP(hipEventRecord(m->ev0, m->stream));
        rc = launch(m, streams_set(m) ? m->d->k_persistent_io : m->d->k_persistent, a);
        if (rc) return rc;
        RF_HIP(hipEventRecord(m->ev1, m->stream));
        RF_HIP(hipEventSynchronize(m->ev1));
        float ms = 0;
        RF_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        m->stats.fused_launches++;
        m->stats.fused_ms += ms;
        m->did_init = true;
        rc = check_flags(m);
    }
    if (rc) return rc;
    q.pos += n;
    if (position) *position = q.pos;
    return RF_OK;
}

int rf_queue_hashes(rf_model* m, uint64_t* out_hashes, uint32_t nstim) {
    if (!m || (!out_hashes && nstim)) return fail(RF_E_ARG, "null argument");
    if (!m->q.open) return fail(RF_E_STATE, "no open queue");
    if (nstim != m->q.nstim) return fail(RF_E_ARG, "queue holds %u stimuli, buffer %u", m->q.nstim, nstim);
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (!nstim) return RF_OK;
    RF_HIP(hipMemcpy(out_hashes, m->out_hash, (size_t)nstim * 8, hipMemcpyDeviceToHost));
    // stimuli that have not run read 0; those of budget 0 that have, the
    // FNV offset basis (as rf_submit_queue returns them)
    std::vector<uint8_t> ran(nstim, 0);
    for (uint32_t k = 0; k < m->q.pos; ++k) ran[m->q.ord[k]] = 1;
    for (uint32_t k = 0; k < nstim; ++k) {
        if (!ran[k]) out_hashes[k] = 0;
        else if (!m->q.budgets[k]) out_hashes[k] = 0xcbf29ce484222325ull;
    }
    return RF_OK;
}

}  // extern "C"

namespace {

constexpr char kQueueMagic[8] = {'R', 'F', 'Q', 'U', 'E', 'U', 'E', 1};

struct queue_header {
    char magic[8];
    uint32_t version, header_bytes;
    uint64_t design_hash;
    uint64_t seed;
    uint32_t nstim, pos;
    uint64_t total_bytes;
    uint64_t reserved[3];
};
static_assert(sizeof(queue_header) == 72, "queue checkpoint header layout");

size_t queue_blob_bytes(uint32_t nstim) { return sizeof(queue_header) + (size_t)nstim * 16; }

}  // namespace

extern "C" {

int rf_queue_save_bytes(rf_model* m, size_t* bytes) {
    if (!m || !bytes) return fail(RF_E_ARG, "null argument");
    if (!m->q.open) return fail(RF_E_STATE, "no open queue");
    *bytes = queue_blob_bytes(m->q.nstim);
    return RF_OK;
}

int rf_queue_save(rf_model* m, void* dst, size_t bytes) {
    if (!m || !dst) return fail(RF_E_ARG, "null argument");
    if (!m->q.open) return fail(RF_E_STATE, "no open queue");
    const uint32_t n = m->q.nstim;
    if (bytes < queue_blob_bytes(n))
        return fail(RF_E_ARG, "queue checkpoint needs %zu bytes, buffer has %zu", queue_blob_bytes(n), bytes);
    {
        const int src = settle(m);
        if (src) return src;
    }
    queue_header h{};
    memcpy(h.magic, kQueueMagic, 8);
    h.version = 1;
    h.header_bytes = sizeof h;
    h.design_hash = m->d->hash;
    h.seed = m->q.seed;
    h.nstim = n;
    h.pos = m->q.pos;
    h.total_bytes = queue_blob_bytes(n);
    uint8_t* o = (uint8_t*)dst;
    memcpy(o, &h, sizeof h);
    o += sizeof h;
    if (n) {
        memcpy(o, m->q.budgets.data(), (size_t)n * 4);
        memcpy(o + (size_t)n * 4, m->q.ord.data(), (size_t)n * 4);
        RF_HIP(hipMemcpy(o + (size_t)n * 8, m->out_hash, (size_t)n * 8, hipMemcpyDeviceToHost));
    }
    return RF_OK;
}

int rf_queue_load(rf_model* m, const void* src, size_t bytes) {
    if (!m || !src) return fail(RF_E_ARG, "null argument");
    queue_header h{};
    if (bytes < sizeof h) return fail(RF_E_ARG, "queue checkpoint shorter than its header");
    memcpy(&h, src, sizeof h);
    if (memcmp(h.magic, kQueueMagic, 8) || h.version != 1 || h.header_bytes != sizeof h)
        return fail(RF_E_ARG, "not a queue checkpoint of this runtime (magic or version)");
    if (h.design_hash != m->d->hash)
        return fail(RF_E_DESIGN, "queue checkpoint of another design (hash %llx, model %llx)",
                    (unsigned long long)h.design_hash, (unsigned long long)m->d->hash);
    if (h.total_bytes != queue_blob_bytes(h.nstim) || bytes < h.total_bytes || h.pos > h.nstim)
        return fail(RF_E_ARG, "queue checkpoint size or position inconsistent");
    const uint8_t* i = (const uint8_t*)src + sizeof h;
    const uint32_t n = h.nstim;
    const uint32_t* budgets = (const uint32_t*)i;
    const uint32_t* order = (const uint32_t*)(i + (size_t)n * 4);
    int rc = rf_queue_open(m, h.seed, budgets, n, order);   // checks the order is a permutation
    if (rc) return rc;
    if (n) RF_HIP(hipMemcpy(m->out_hash, i + (size_t)n * 8, (size_t)n * 8, hipMemcpyHostToDevice));
    m->q.pos = h.pos;
    return RF_OK;
}

int rf_set_mem_images(rf_model* m, uint32_t mi, uint32_t nstim, uint32_t nelem, const void* src) {
    if (!m) return fail(RF_E_ARG, "null model");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (mi >= m->d->nmems) return fail(RF_E_ARG, "memory %u out of range", mi);
    const rf_mem& md = m->d->mems[mi];
    if (nelem > md.depth) return fail(RF_E_ARG, "image larger than the memory");
    if (m->image_bufs[mi]) {
        RF_HIP(hipFree(m->image_bufs[mi]));
        m->image_bufs[mi] = nullptr;
    }
    uint64_t desc[2] = {0, 0};
    if (src && nstim && nelem) {
        const size_t bytes = (size_t)nstim * nelem * md.words * kClassBytes[md.cls];
        RF_HIP(hipMalloc(&m->image_bufs[mi], bytes));
        RF_HIP(hipMemcpyAsync(m->image_bufs[mi], src, bytes, hipMemcpyHostToDevice, m->stream));
        desc[0] = (uint64_t)(uintptr_t)m->image_bufs[mi];
        desc[1] = ((uint64_t)nstim << 32) | nelem;
    }
    RF_HIP(hipMemcpyAsync(m->images + (size_t)mi * 16, desc, 16, hipMemcpyHostToDevice, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    bool any = false;
    for (void* p : m->image_bufs) any = any || p;
    m->base.images = any ? m->images : nullptr;
    return RF_OK;
}

// Queue buffers for nstim stimuli and the queue order (the caller's, or
// longest budget first).
static int queue_setup(rf_model* m, const uint32_t* budgets, uint32_t nstim, const uint32_t* order,
                std::vector<uint32_t>& ord) {
    if (nstim > m->qcap) {
        if (m->budget) (void)hipFree(m->budget);
        if (m->order) (void)hipFree(m->order);
        if (m->out_hash) (void)hipFree(m->out_hash);
        m->budget = nullptr;
        m->order = nullptr;
        m->out_hash = nullptr;
        m->qcap = 0;
        RF_HIP(hipMalloc(&m->budget, (size_t)nstim * 4));
        RF_HIP(hipMalloc(&m->order, (size_t)nstim * 4));
        RF_HIP(hipMalloc(&m->out_hash, (size_t)nstim * 8));
        m->qcap = nstim;
    }
    ord.resize(nstim);
    if (order) {
        std::vector<char> seen(nstim, 0);
        for (uint32_t k = 0; k < nstim; ++k) {
            if (order[k] >= nstim || seen[order[k]])
                return fail(RF_E_ARG, "order is not a permutation of [0, %u)", nstim);
            seen[order[k]] = 1;
            ord[k] = order[k];
        }
    } else {
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(),
                         [&](uint32_t x, uint32_t y) { return budgets[x] > budgets[y]; });
    }
    return RF_OK;
}

// The persistent launch of a batched workload over device queue buffers
// (budget, order, out_hash), the host order ``ord`` already checked.
static int run_queue(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim,
                     const uint32_t* ord, uint32_t* dbudget, uint32_t* dorder, uint64_t* dout,
                     uint64_t* out_hashes) {
    RF_HIP(hipMemcpyAsync(dbudget, budgets, (size_t)nstim * 4, hipMemcpyHostToDevice, m->stream));
    RF_HIP(hipMemcpyAsync(dorder, ord, (size_t)nstim * 4, hipMemcpyHostToDevice, m->stream));
    uint32_t f[RF_FLAG_WORDS] = {0};
    f[RF_FLAG_QUEUE] = m->nlanes;
    RF_HIP(hipMemcpyAsync(m->flags, f, sizeof f, hipMemcpyHostToDevice, m->stream));
    rf_args a = args_for(m, RF_PHASE_EVAL);
    a.seed = seed;
    a.budget = dbudget;
    a.order = dorder;
    a.out_hash = dout;
    a.nstim = nstim;
    RF_HIP(hipEventRecord(m->ev0, m->stream));
    int rc = launch(m, streams_set(m) ? m->d->k_persistent_io : m->d->k_persistent, a);
    if (rc) return rc;
    RF_HIP(hipEventRecord(m->ev1, m->stream));
    RF_HIP(hipEventSynchronize(m->ev1));
    float ms = 0;
    RF_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->stats.fused_launches++;
    m->stats.fused_ms += ms;
    RF_HIP(hipMemcpyAsync(out_hashes, dout, (size_t)nstim * 8, hipMemcpyDeviceToHost, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    m->did_init = true;
    return check_flags(m);
}

int rf_submit_queue(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim,
                    const uint32_t* order, uint64_t* out_hashes) {
    if (!m || !budgets || !out_hashes) return fail(RF_E_ARG, "null argument");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (!nstim) return RF_OK;
    int qrc = queue_stream_check(m, budgets, nstim);
    if (qrc) return qrc;
    std::vector<uint32_t> ord;
    qrc = queue_setup(m, budgets, nstim, order, ord);
    if (qrc) return qrc;
    return run_queue(m, seed, budgets, nstim, ord.data(), m->budget, m->order, m->out_hash, out_hashes);
}

int rf_submit_queue_static(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim,
                           const uint32_t* order, uint64_t* out_hashes) {
    if (!m || !budgets || !out_hashes) return fail(RF_E_ARG, "null argument");
    if (nstim > RF_QUEUE_STATIC_MAX)
        return fail(RF_E_ARG, "rf_submit_queue_static: %u stimuli exceed its fixed queue of %u "
                    "(use rf_submit_queue)", nstim, (uint32_t)RF_QUEUE_STATIC_MAX);
    if (!m->queue_scratch)
        return fail(RF_E_HIP, "rf_submit_queue_static: no device address for the queue buffers "
                    "(hipGetSymbolAddress: %s)", hipGetErrorString((hipError_t)m->queue_scratch_err));
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (!nstim) return RF_OK;
    {
        const int qrc = queue_stream_check(m, budgets, nstim);
        if (qrc) return qrc;
    }
    std::lock_guard<std::mutex> lk(g_static_order_mu);
    uint32_t* ord = g_static_order;
    if (order) {
        memset(g_static_seen, 0, ((size_t)nstim + 63) / 64 * 8);
        for (uint32_t k = 0; k < nstim; ++k) {
            const uint32_t x = order[k];
            if (x >= nstim || (g_static_seen[x >> 6] >> (x & 63)) & 1)
                return fail(RF_E_ARG, "order is not a permutation of [0, %u)", nstim);
            g_static_seen[x >> 6] |= 1ull << (x & 63);
            ord[k] = x;
        }
    } else {
        // longest budget first; ties by stimulus id (the order the
        // allocating form's stable sort gives), sorted in place
        for (uint32_t k = 0; k < nstim; ++k) ord[k] = k;
        std::sort(ord, ord + nstim, [&](uint32_t x, uint32_t y) {
            return budgets[x] != budgets[y] ? budgets[x] > budgets[y] : x < y;
        });
    }
    uint8_t* base = m->queue_scratch;
    return run_queue(m, seed, budgets, nstim, ord, (uint32_t*)(base + kScratchBudgetOff),
                     (uint32_t*)(base + kScratchOrderOff), (uint64_t*)(base + kScratchHashOff),
                     out_hashes);
}

// The bit-plane form of a batched workload over queue positions
// ord[0, n): rounds of n_lanes stimuli, every round from the reset state;
// each stimulus's hash goes to the device out_hash[stimulus id] (nstim: the
// whole queue's size, for the kernel's budget lookups).
static int bitplane_rounds(rf_model* m, uint64_t seed, const uint32_t* budgets, const uint32_t* ord,
                           uint32_t n, uint32_t nstim) {
    int rc = RF_OK;
    const uint32_t nl = m->nlanes;
    std::vector<uint64_t> h(m->npad, 0xcbf29ce484222325ull);
    std::vector<uint32_t> sid(m->npad, 0);
    std::vector<uint8_t> done(m->npad, 1);
    for (uint32_t r0 = 0; r0 < n; r0 += nl) {
        const uint32_t cnt = std::min(nl, n - r0);
        uint32_t ncyc = 0;
        for (uint32_t k = 0; k < nl; ++k) {
            done[k] = k < cnt ? 0 : 1;
            sid[k] = k < cnt ? ord[r0 + k] : 0;
            if (k < cnt) ncyc = std::max(ncyc, budgets[ord[r0 + k]]);
        }
        // every lane starts from the reset state: zeroed arrays and memories,
        // then the initial blocks
        for (int c = 0; c < 4; ++c)
            if (m->d->slots[c])
                RF_HIP(hipMemsetAsync(m->cls[c], 0, m->d->slots[c] * m->npad * kClassBytes[c], m->stream));
        if (m->mem) RF_HIP(hipMemsetAsync(m->mem, 0, m->d->mem_row_bytes * m->npad, m->stream));
        m->did_init = false;
        rc = do_init(m);
        if (rc) return rc;
        RF_HIP(hipMemcpyAsync(m->hash, h.data(), m->npad * 8, hipMemcpyHostToDevice, m->stream));
        RF_HIP(hipMemcpyAsync(m->stim, sid.data(), m->npad * 4, hipMemcpyHostToDevice, m->stream));
        RF_HIP(hipMemcpyAsync(m->done, done.data(), m->npad, hipMemcpyHostToDevice, m->stream));
        RF_HIP(hipMemsetAsync(m->cyc, 0, m->npad * 4, m->stream));
        if (!ncyc) continue;
        rf_bp_params p = m->bp;
        rf_bp_rounds(m->d->bitplane, m->cfg.bitplane_threads, &p);
        p.a = args_for(m, RF_PHASE_EVAL);
        p.a.seed = seed;
        p.a.ncycles = ncyc;
        p.a.budget = m->budget;
        p.a.out_hash = m->out_hash;
        p.a.nstim = nstim;
        RF_HIP(hipMemsetAsync(m->bp_progress, 0, sizeof(unsigned long long), m->stream));
        RF_HIP(hipEventRecord(m->ev0, m->stream));
        RF_HIP(rf_bp_launch(p, m->cfg.bitplane_threads, m->stream));
        RF_HIP(hipEventRecord(m->ev1, m->stream));
        RF_HIP(hipEventSynchronize(m->ev1));
        float ms = 0;
        RF_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        m->stats.bitplane_launches++;
        m->stats.bitplane_ms += ms;
    }
    std::vector<uint8_t> undone(m->npad, 0);
    RF_HIP(hipMemcpyAsync(m->done, undone.data(), m->npad, hipMemcpyHostToDevice, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    return check_flags(m);
}

int rf_submit_queue_bitplane(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim,
                             const uint32_t* order, uint64_t* out_hashes) {
    if (!m || !budgets || !out_hashes) return fail(RF_E_ARG, "null argument");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (!m->bp_prog) return fail(RF_E_STATE, "design %s has no bit-plane program", m->d->name);
    if (m->base.images)
        return fail(RF_E_STATE, "memory images are loaded by the per-lane batched form only");
    if (!nstim) return RF_OK;
    int rc = queue_stream_check(m, budgets, nstim);
    if (rc) return rc;
    std::vector<uint32_t> ord;
    rc = queue_setup(m, budgets, nstim, order, ord);
    if (rc) return rc;
    RF_HIP(hipMemsetAsync(m->flags, 0, RF_FLAG_WORDS * 4, m->stream));
    RF_HIP(hipMemcpyAsync(m->budget, budgets, (size_t)nstim * 4, hipMemcpyHostToDevice, m->stream));
    RF_HIP(hipMemsetAsync(m->out_hash, 0, (size_t)nstim * 8, m->stream));
    rc = bitplane_rounds(m, seed, budgets, ord.data(), nstim, nstim);
    if (rc) return rc;
    // stimuli of budget 0 keep the FNV offset basis
    std::vector<uint64_t> oh(nstim);
    RF_HIP(hipMemcpy(oh.data(), m->out_hash, (size_t)nstim * 8, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < nstim; ++k) out_hashes[k] = budgets[k] ? oh[k] : 0xcbf29ce484222325ull;
    return RF_OK;
}

// ---------------------------------------------------------------- checkpoints
//
// A checkpoint is one byte blob: a header naming the design and the layout,
// then every device array of the model in a fixed order. Both ends of a copy
// may be host or device memory (hipMemcpyDefault).

}  // extern "C"

namespace {

constexpr char kCkptMagic[8] = {'R', 'F', 'C', 'K', 'P', 'T', 0, 1};
constexpr uint32_t kCkptVersion = 1;

struct ckpt_header {
    char magic[8];
    uint32_t version, header_bytes;
    uint64_t design_hash;
    uint32_t nlanes, did_init;
    uint64_t npad;
    uint64_t slots[4];
    uint64_t mem_row_bytes;
    uint64_t seed;
    uint64_t total_bytes;
    uint64_t reserved[4];
};
static_assert(sizeof(ckpt_header) == 128, "checkpoint header is 128 bytes");

struct ckpt_section {
    void* dev;
    size_t bytes;
};

std::vector<ckpt_section> ckpt_sections(const rf_model* m) {
    std::vector<ckpt_section> out;
    for (int c = 0; c < 4; ++c) out.push_back({m->cls[c], m->d->slots[c] * m->npad * kClassBytes[c]});
    out.push_back({m->mem, m->d->mem_row_bytes * m->npad});
    out.push_back({m->change, m->npad * 4});
    out.push_back({m->done, m->npad});
    out.push_back({m->hash, m->npad * 8});
    out.push_back({m->cyc, m->npad * 4});
    out.push_back({m->stim, m->npad * 4});
    out.push_back({m->flags, RF_FLAG_WORDS * 4});
    return out;
}

size_t ckpt_total(const rf_model* m) {
    size_t n = sizeof(ckpt_header);
    for (const auto& sct : ckpt_sections(m)) n += sct.bytes;
    return n;
}

}  // namespace

extern "C" {

int rf_checkpoint_bytes(rf_model* m, size_t* bytes) {
    if (!m || !bytes) return fail(RF_E_ARG, "null argument");
    *bytes = ckpt_total(m);
    return RF_OK;
}

int rf_checkpoint_save(rf_model* m, void* dst, size_t bytes) {
    if (!m || !dst) return fail(RF_E_ARG, "null model or buffer");
    {
        const int src = settle(m);
        if (src) return src;
    }
    const size_t total = ckpt_total(m);
    if (bytes < total) return fail(RF_E_ARG, "checkpoint buffer too small (%zu < %zu)", bytes, total);
    ckpt_header h{};
    memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.header_bytes = sizeof h;
    h.design_hash = m->d->hash;
    h.nlanes = m->nlanes;
    h.did_init = m->did_init ? 1u : 0u;
    h.npad = m->npad;
    for (int c = 0; c < 4; ++c) h.slots[c] = m->d->slots[c];
    h.mem_row_bytes = m->d->mem_row_bytes;
    h.seed = m->base.seed;
    h.total_bytes = total;
    uint8_t* o = (uint8_t*)dst;
    RF_HIP(hipMemcpyAsync(o, &h, sizeof h, hipMemcpyDefault, m->stream));
    o += sizeof h;
    for (const auto& sct : ckpt_sections(m)) {
        if (sct.bytes) RF_HIP(hipMemcpyAsync(o, sct.dev, sct.bytes, hipMemcpyDefault, m->stream));
        o += sct.bytes;
    }
    RF_HIP(hipStreamSynchronize(m->stream));
    return RF_OK;
}

int rf_checkpoint_load(rf_model* m, const void* src, size_t bytes) {
    if (!m || !src) return fail(RF_E_ARG, "null model or buffer");
    {
        const int rc = settle(m);
        if (rc) return rc;
    }
    ckpt_header h{};
    if (bytes < sizeof h) return fail(RF_E_ARG, "checkpoint shorter than its header");
    RF_HIP(hipMemcpyAsync(&h, src, sizeof h, hipMemcpyDefault, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    if (memcmp(h.magic, kCkptMagic, 8) || h.version != kCkptVersion || h.header_bytes != sizeof h)
        return fail(RF_E_ARG, "not a checkpoint of this runtime (magic or version)");
    if (h.design_hash != m->d->hash)
        return fail(RF_E_DESIGN, "checkpoint of another design (hash %llx, model %llx)",
                    (unsigned long long)h.design_hash, (unsigned long long)m->d->hash);
    if (h.nlanes != m->nlanes || h.npad != m->npad)
        return fail(RF_E_ARG, "checkpoint of %u lanes, model has %u", h.nlanes, m->nlanes);
    for (int c = 0; c < 4; ++c)
        if (h.slots[c] != m->d->slots[c]) return fail(RF_E_DESIGN, "checkpoint layout differs");
    if (h.mem_row_bytes != m->d->mem_row_bytes) return fail(RF_E_DESIGN, "checkpoint layout differs");
    const size_t total = ckpt_total(m);
    if (h.total_bytes != total || bytes < total)
        return fail(RF_E_ARG, "checkpoint size %zu, expected %zu", bytes, total);
    const uint8_t* i = (const uint8_t*)src + sizeof h;
    for (const auto& sct : ckpt_sections(m)) {
        if (sct.bytes) RF_HIP(hipMemcpyAsync(sct.dev, i, sct.bytes, hipMemcpyDefault, m->stream));
        i += sct.bytes;
    }
    m->did_init = h.did_init != 0;
    if (h.seed != m->base.seed) m->graphs_stale = true;
    m->base.seed = h.seed;
    return RF_OK;
}

int rf_lane_state(rf_model* m, uint32_t lane, void* dst, size_t* bytes) {
    if (!m || !bytes) return fail(RF_E_ARG, "null argument");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (lane >= m->nlanes) return fail(RF_E_ARG, "lane %u out of range", lane);
    size_t total = 0;
    for (int c = 0; c < 4; ++c) total += m->d->slots[c] * kClassBytes[c];
    total += m->d->mem_row_bytes;
    if (!dst) {
        *bytes = total;
        return RF_OK;
    }
    if (*bytes < total) return fail(RF_E_ARG, "buffer too small (%zu < %zu)", *bytes, total);
    uint8_t* o = (uint8_t*)dst;
    for (int c = 0; c < 4; ++c) {
        const size_t es = kClassBytes[c];
        if (!m->d->slots[c]) continue;
        // one strided copy: slot k of this lane at (k * npad + lane) * es
        RF_HIP(hipMemcpy2DAsync(o, es, (uint8_t*)m->cls[c] + lane * es, m->npad * es, es,
                                m->d->slots[c], hipMemcpyDeviceToHost, m->stream));
        o += m->d->slots[c] * es;
    }
    for (uint32_t k = 0; k < m->d->nmems; ++k) {
        const rf_mem& md = m->d->mems[k];
        if (md.layout != RF_MEM_LANE_MAJOR) continue;
        RF_HIP(hipMemcpyAsync(o + md.loc, m->mem + md.loc * m->npad + (size_t)lane * md.row_bytes,
                              md.row_bytes, hipMemcpyDeviceToHost, m->stream));
    }
    RF_HIP(hipStreamSynchronize(m->stream));
    *bytes = total;
    return RF_OK;
}

int rf_trace_lane(rf_model* m, uint32_t lane, uint32_t ncycles, uint32_t mode, void* dst, size_t bytes) {
    if (!m || !dst) return fail(RF_E_ARG, "null model or buffer");
    {
        const int src = settle(m);
        if (src) return src;
    }
    if (lane >= m->nlanes) return fail(RF_E_ARG, "lane %u out of range", lane);
    if (mode > RF_MODE_BITPLANE) return fail(RF_E_ARG, "unknown run mode %u", mode);
    size_t sb = m->d->mem_row_bytes;
    for (int c = 0; c < 4; ++c) sb += m->d->slots[c] * kClassBytes[c];
    if (bytes < (size_t)ncycles * sb)
        return fail(RF_E_ARG, "trace buffer too small (%zu < %zu)", bytes, (size_t)ncycles * sb);
    if (!ncycles) return RF_OK;
    uint8_t* tb = nullptr;
    RF_HIP(hipMalloc(&tb, (size_t)ncycles * sb));
    int rc = RF_OK;
    for (uint32_t c = 0; c < ncycles && rc == RF_OK; ++c) {
        // one launch (or graph replay) per cycle, then the lane's slots
        // copied out of the lane-minor arrays on the same stream
        const uint32_t chunk = m->cfg.fused_chunk;
        m->cfg.fused_chunk = 1;
        rc = rf_run_cycles(m, 1, mode);
        m->cfg.fused_chunk = chunk;
        uint8_t* o = tb + (size_t)c * sb;
        for (int k = 0; k < 4 && rc == RF_OK; ++k) {
            const size_t es = kClassBytes[k];
            if (!m->d->slots[k]) continue;
            const hipError_t e = hipMemcpy2DAsync(o, es, (uint8_t*)m->cls[k] + lane * es, m->npad * es, es,
                                                  m->d->slots[k], hipMemcpyDeviceToDevice, m->stream);
            if (e != hipSuccess) rc = fail(RF_E_HIP, "trace copy: %s", hipGetErrorString(e));
            o += m->d->slots[k] * es;
        }
        for (uint32_t k = 0; k < m->d->nmems && rc == RF_OK; ++k) {
            const rf_mem& md = m->d->mems[k];
            if (md.layout != RF_MEM_LANE_MAJOR) continue;
            const hipError_t e = hipMemcpyAsync(o + md.loc, m->mem + md.loc * m->npad + (size_t)lane * md.row_bytes,
                                                md.row_bytes, hipMemcpyDeviceToDevice, m->stream);
            if (e != hipSuccess) rc = fail(RF_E_HIP, "trace copy: %s", hipGetErrorString(e));
        }
    }
    if (rc == RF_OK) rc = settle(m);
    if (rc == RF_OK) {
        const hipError_t e = hipMemcpy(dst, tb, (size_t)ncycles * sb, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(RF_E_HIP, "trace download: %s", hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(tb);
    return rc;
}

int rf_profile_mtasks(rf_model* m, uint32_t reps, float* ms_out) {
    if (!m || !ms_out || !reps) return fail(RF_E_ARG, "bad argument");
    {
        const int src = settle(m);
        if (src) return src;
    }
    int rc = do_init(m);
    if (rc) return rc;
    // every lane requests a pass, so kernels that skip settled lanes do work
    {
        std::vector<uint32_t> ones(m->npad, 1u);
        RF_HIP(hipMemcpyAsync(m->change, ones.data(), m->npad * 4, hipMemcpyHostToDevice, m->stream));
        RF_HIP(hipStreamSynchronize(m->stream));
    }
    const rf_args a = args_for(m, RF_PHASE_EVAL);
    for (uint32_t k = 0; k < m->d->nmtasks; ++k) {
        // one warm-up launch, then ``reps`` timed launches back to back
        rc = launch(m, m->d->mtasks[k].kernel, a);
        if (rc) return rc;
        RF_HIP(hipEventRecord(m->ev0, m->stream));
        for (uint32_t r = 0; r < reps; ++r) {
            rc = launch(m, m->d->mtasks[k].kernel, a);
            if (rc) return rc;
        }
        RF_HIP(hipEventRecord(m->ev1, m->stream));
        RF_HIP(hipEventSynchronize(m->ev1));
        float ms = 0;
        RF_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        ms_out[k] = ms / (float)reps;
    }
    return RF_OK;
}

int rf_sync(rf_model* m) {
    if (!m) return fail(RF_E_ARG, "null model");
    {
        const int src = settle(m);
        if (src) return src;
    }
    RF_HIP(hipStreamSynchronize(m->stream));
    return RF_OK;
}

int rf_flags(rf_model* m, uint32_t* out, uint32_t n) {
    if (!m || !out) return fail(RF_E_ARG, "null argument");
    {
        const int src = settle(m, false);
        if (src) return src;
    }
    if (n > RF_FLAG_WORDS) n = RF_FLAG_WORDS;
    RF_HIP(hipMemcpyAsync(out, m->flags, n * 4, hipMemcpyDeviceToHost, m->stream));
    RF_HIP(hipStreamSynchronize(m->stream));
    return RF_OK;
}

int rf_get_stats(rf_model* m, rf_stats* out) {
    if (!m || !out) return fail(RF_E_ARG, "null argument");
    {
        const int src = settle(m, false);
        if (src) return src;
    }
    *out = m->stats;
    return RF_OK;
}

int rf_reset_stats(rf_model* m) {
    if (!m) return fail(RF_E_ARG, "null model");
    {
        const int src = settle(m, false);
        if (src) return src;
    }
    m->stats = rf_stats{};
    return RF_OK;
}

}  // extern "C"
