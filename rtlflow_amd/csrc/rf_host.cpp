// rf_host.cpp -- driver for the one-lane host build of a generated design.
//
// The emitter can write every kernel as a host function with the identical
// body (``void k(rf_args a, uint32_t lane)``); this driver replays them lane
// by lane with the same protocol the GPU runtime uses. It exists for tests:
// the survey's "one-lane host build of the emitted model", and a race test
// that runs the mtasks of each eval in a random topological order (any order
// the hipGraph may choose must give the same hashes). It is never used on the
// GPU path.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "rf_design.h"

namespace {

typedef void (*host_kernel)(rf_args, uint32_t);

uint32_t g_state_lane = ~0u;
uint8_t* g_state_out = nullptr;
rf_stream g_in{}, g_out{};  // input stream / output capture of later runs

struct HostModel {
    const rf_design* d;
    uint32_t n;
    uint64_t npad;
    std::vector<uint8_t> c, mem, done;
    std::vector<uint16_t> s;
    std::vector<uint32_t> i, change, cyc, stim, flags;
    std::vector<uint64_t> q, hash;
    rf_args a{};

    HostModel(const rf_design* d_, uint32_t n_) : d(d_), n(n_) {
        npad = ((uint64_t)n + 63) / 64 * 64;
        c.assign(std::max<uint64_t>(1, d->slots[0]) * npad, 0);
        s.assign(std::max<uint64_t>(1, d->slots[1]) * npad, 0);
        i.assign(std::max<uint64_t>(1, d->slots[2]) * npad, 0);
        q.assign(std::max<uint64_t>(1, d->slots[3]) * npad, 0);
        mem.assign(std::max<uint64_t>(8, d->mem_row_bytes) * npad, 0);
        done.assign(npad, 0);
        change.assign(npad, 1);
        cyc.assign(npad, 0);
        stim.assign(npad, 0);
        flags.assign(RF_FLAG_WORDS, 0);
        hash.assign(npad, 0xcbf29ce484222325ull);
        a.c = c.data();
        a.s = s.data();
        a.i = i.data();
        a.q = q.data();
        a.mem = mem.data();
        a.change = change.data();
        a.done = done.data();
        a.hash = hash.data();
        a.cyc = cyc.data();
        a.stim = stim.data();
        a.flags = flags.data();
        a.npad = npad;
        a.nlanes = n;
        a.converge_limit = 100;
        a.in = g_in;
        a.out = g_out;
    }
    void all(const void* k, uint32_t phase) {
        rf_args x = a;
        x.phase = phase;
        host_kernel f = (host_kernel)k;
        for (uint32_t l = 0; l < n; ++l) f(x, l);
    }
};

// One eval with the host-driven settle loop; ``order`` lists mtask indices in
// the execution order to use.
int eval(HostModel& m, const std::vector<uint32_t>& order, uint32_t final_phase) {
    const rf_design* d = m.d;
    for (uint32_t it = 0;; ++it) {
        m.flags[RF_FLAG_CHANGE] = 0;
        for (uint32_t k : order) m.all(d->mtasks[k].kernel, RF_PHASE_EVAL);
        if (!d->nchange) break;
        m.all(d->k_eval_end, RF_PHASE_EVAL);
        if (!m.flags[RF_FLAG_CHANGE]) break;
        if (it + 1 >= m.a.converge_limit) return -4;
    }
    m.all(d->k_eval_end, final_phase);
    return 0;
}

// One eval as the device's cycle graph runs it: ``unroll`` passes of the
// mtask DAG (each masked by the lanes' change requests and closed by an
// eval-end), then the settle-tail kernel for lanes still changing, then the
// eval-end that does the phase work.
int eval_graph_form(HostModel& m, const std::vector<uint32_t>& order, uint32_t final_phase,
                    uint32_t unroll) {
    const rf_design* d = m.d;
    if (!d->nchange || !d->k_settle) {
        for (uint32_t k : order) m.all(d->mtasks[k].kernel, RF_PHASE_EVAL);
        m.all(d->k_eval_end, final_phase);
        return 0;
    }
    for (uint32_t it = 0; it < unroll; ++it) {
        for (uint32_t k : order) m.all(d->mtasks[k].kernel, RF_PHASE_EVAL);
        m.all(d->k_eval_end, RF_PHASE_EVAL);
    }
    rf_args x = m.a;
    x.ncycles = unroll;
    host_kernel f = (host_kernel)d->k_settle;
    for (uint32_t l = 0; l < m.n; ++l) f(x, l);
    m.all(d->k_eval_end, final_phase);
    return 0;
}

std::vector<uint32_t> topo_order(const rf_design* d, std::mt19937_64* rng) {
    const uint32_t n = d->nmtasks;
    std::vector<uint32_t> indeg(n, 0);
    std::vector<std::vector<uint32_t>> succ(n);
    for (uint32_t k = 0; k < n; ++k)
        for (uint32_t j = 0; j < d->mtasks[k].npreds; ++j) {
            succ[d->mtasks[k].preds[j]].push_back(k);
            indeg[k]++;
        }
    std::vector<uint32_t> ready, out;
    for (uint32_t k = 0; k < n; ++k)
        if (!indeg[k]) ready.push_back(k);
    while (!ready.empty()) {
        size_t pick = 0;
        if (rng) pick = (*rng)() % ready.size();
        const uint32_t k = ready[pick];
        ready.erase(ready.begin() + pick);
        out.push_back(k);
        for (uint32_t s : succ[k])
            if (--indeg[s] == 0) ready.push_back(s);
    }
    return out;
}

}  // namespace

#define RF_EXPORT __attribute__((visibility("default")))

extern "C" {

// Copy per-stimulus memory images (same table layout as rf_set_mem_images:
// {base pointer, (nstim << 32) | nelem} per memory) into each lane.
static void load_images(HostModel& m, const uint64_t* images, uint32_t stim0) {
    if (!images) return;
    static const size_t es[4] = {1, 2, 4, 8};
    for (uint32_t j = 0; j < m.d->nmems; ++j) {
        const rf_mem& md = m.d->mems[j];
        const uint8_t* src = (const uint8_t*)images[2 * j];
        const uint32_t ne = (uint32_t)images[2 * j + 1];
        const uint32_t ns = (uint32_t)(images[2 * j + 1] >> 32);
        if (!src) continue;
        const size_t esz = es[md.cls];
        uint8_t* base[4] = {m.c.data(), (uint8_t*)m.s.data(), (uint8_t*)m.i.data(), (uint8_t*)m.q.data()};
        for (uint32_t l = 0; l < m.n; ++l) {
            const uint32_t sid = stim0 + l;
            if (sid >= ns) continue;
            const uint8_t* img = src + (size_t)sid * ne * md.words * esz;
            for (uint32_t k = 0; k < ne * md.words; ++k) {
                uint8_t* dst = md.layout == RF_MEM_LANE_MAJOR
                                   ? m.mem.data() + md.loc * m.npad + (size_t)l * md.row_bytes + k * esz
                                   : base[md.cls] + ((md.loc + k) * m.npad + l) * esz;
                memcpy(dst, img + k * esz, esz);
            }
        }
    }
}

// mode 0: mtasks in table order; 1: fused kernel; 2: random topological order
// per eval (seeded by order_seed); 3: the device cycle graph's schedule with
// two captured settle passes and the settle-tail kernel. images: optional
// per-stimulus memory images.
RF_EXPORT int rf_host_protocol(const rf_design* d, uint32_t nlanes, uint64_t seed, uint32_t stim0,
                               uint32_t ncycles, uint32_t mode, uint64_t order_seed, uint64_t* hashes,
                               const uint64_t* images) {
    HostModel m(d, nlanes);
    m.a.seed = seed;
    for (uint32_t l = 0; l < nlanes; ++l) m.stim[l] = stim0 + l;
    load_images(m, images, stim0);
    m.all(d->k_init, RF_PHASE_EVAL);
    std::mt19937_64 rng(order_seed);
    if (mode == 1) {
        rf_args x = m.a;
        x.ncycles = ncycles;
        host_kernel f = (host_kernel)(x.in.data || x.out.data ? d->k_fused_io : d->k_fused);
        for (uint32_t l = 0; l < nlanes; ++l) f(x, l);
    } else {
        for (uint32_t c = 0; c < ncycles; ++c) {
            m.all(d->k_stim, RF_PHASE_EVAL);
            auto o1 = topo_order(d, mode == 2 ? &rng : nullptr);
            if (mode == 3 ? eval_graph_form(m, o1, RF_PHASE_RISE, 2) : eval(m, o1, RF_PHASE_RISE))
                return -4;
            auto o2 = topo_order(d, mode == 2 ? &rng : nullptr);
            if (mode == 3 ? eval_graph_form(m, o2, RF_PHASE_HASH, 2) : eval(m, o2, RF_PHASE_HASH))
                return -4;
        }
    }
    if (m.flags[RF_FLAG_UNCONVERGED]) return -4;
    if (m.flags[RF_FLAG_STREAM]) return -1;
    memcpy(hashes, m.hash.data(), nlanes * 8ull);
    if (g_state_lane < nlanes && g_state_out) {
        // raw state of one lane in the rf_lane_state order: C, S, I, Q slots,
        // then the lane-major memory region
        uint8_t* o = g_state_out;
        for (uint64_t k = 0; k < d->slots[0]; ++k) *o++ = m.c[k * m.npad + g_state_lane];
        for (uint64_t k = 0; k < d->slots[1]; ++k, o += 2) memcpy(o, &m.s[k * m.npad + g_state_lane], 2);
        for (uint64_t k = 0; k < d->slots[2]; ++k, o += 4) memcpy(o, &m.i[k * m.npad + g_state_lane], 4);
        for (uint64_t k = 0; k < d->slots[3]; ++k, o += 8) memcpy(o, &m.q[k * m.npad + g_state_lane], 8);
        for (uint32_t j = 0; j < d->nmems; ++j) {
            const rf_mem& md = d->mems[j];
            if (md.layout != RF_MEM_LANE_MAJOR) continue;
            memcpy(o + md.loc, m.mem.data() + md.loc * m.npad + (size_t)g_state_lane * md.row_bytes,
                   md.row_bytes);
        }
    }
    return 0;
}

// Select a lane whose raw state rf_host_protocol copies out after the run.
RF_EXPORT void rf_host_capture(uint32_t lane, uint8_t* out) {
    g_state_lane = lane;
    g_state_out = out;
}

// Input stream and output capture (host memory here) for the runs that
// follow; NULL: the generator / no capture.
RF_EXPORT void rf_host_streams(const rf_stream* in, const rf_stream* out) {
    g_in = in ? *in : rf_stream{};
    g_out = out ? *out : rf_stream{};
}

RF_EXPORT int rf_host_queue(const rf_design* d, uint32_t nlanes, uint64_t seed, const uint32_t* budgets,
                            uint32_t nstim, const uint32_t* order, uint64_t* out,
                            const uint64_t* images) {
    HostModel m(d, nlanes);
    m.a.images = (const uint8_t*)images;
    std::vector<uint32_t> ord(nstim);
    for (uint32_t k = 0; k < nstim; ++k) ord[k] = order ? order[k] : k;
    m.flags[RF_FLAG_QUEUE] = nlanes;
    rf_args x = m.a;
    x.seed = seed;
    x.budget = budgets;
    x.order = ord.data();
    x.out_hash = out;
    x.nstim = nstim;
    host_kernel f = (host_kernel)(x.in.data || x.out.data ? d->k_persistent_io : d->k_persistent);
    for (uint32_t l = 0; l < nlanes; ++l) f(x, l);
    return m.flags[RF_FLAG_UNCONVERGED] ? -4 : m.flags[RF_FLAG_STREAM] ? -1 : 0;
}

}  // extern "C"
