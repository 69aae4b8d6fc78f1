/* rf_verilated.h -- C ABI of the MI355X batch-stimulus simulation runtime
 * (librf_verilated.so).
 *
 * Replaces the reference's generated model runtime, the ``RF::RTLflow`` class
 * (src/V3EmitC.cpp:5246-5500) plus the ``Vtop`` model surface
 * (src/V3EmitC.cpp:3510-3792), with a plain-C ABI that any host (C++,
 * ctypes, cgo, JNI) can bind. A model owns all device memory of one design
 * instantiated for ``n_lanes`` stimuli; one host thread per model.
 *
 * Errors: every entry point returns 0 on success or a negative RF_E* code
 * and records a message readable with rf_last_error() (thread-local). The
 * reference throws std::runtime_error on CUDA failures
 * (src/V3EmitC.cpp:5327-5333) and aborts through VL_FATAL_MT on
 * non-convergence (src/V3EmitC.cpp:5455-5462); neither crosses this ABI.
 */
#ifndef RF_VERILATED_H
#define RF_VERILATED_H
#include <stddef.h>
#include <stdint.h>

#include "rf_design.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RF_OK = 0,
    RF_E_ARG = -1,         /* bad argument (null, out of range, size mismatch) */
    RF_E_HIP = -2,         /* HIP runtime error (allocation, launch, copy) */
    RF_E_DESIGN = -3,      /* design table ABI mismatch */
    RF_E_CONVERGE = -4,    /* settle loop hit the converge limit */
    RF_E_STATE = -5        /* call not valid in the model's current state */
};

/* run_cycles modes */
enum { RF_MODE_GRAPH = 0, RF_MODE_FUSED = 1, RF_MODE_BITPLANE = 2 };

typedef struct rf_config {
    uint32_t settle_unroll;   /* graph form: settle iterations captured per eval (>= 1) */
    uint32_t converge_limit;  /* settle iterations before RF_E_CONVERGE (reference: 100) */
    uint32_t fused_chunk;     /* fused / bit-plane forms: cycles per kernel launch */
    uint32_t bitplane_threads; /* bit-plane form: threads per 64-lane workgroup
                                  (128, 256, 512, 1024; 0 = 512) */
    uint32_t lane_pad;        /* SoA padding: extra lanes in every slot's row
                                 beyond n_lanes rounded up to the block
                                 (a multiple of 64; 0 = none) */
    uint32_t bitplane_slots;  /* bit-plane form, rf_run_cycles: 64-lane
                                 groups one launch runs at once. 0 = the
                                 workgroups the GPU holds at once, and groups
                                 beyond them rotate through launches of one
                                 full round each when that is estimated
                                 faster than a launch of several rounds;
                                 N > 0 = rotate through windows of at most N
                                 groups whenever there are more than N */
    uint32_t fused_multiwave; /* RF_MODE_FUSED in rf_run_cycles: 1 = the
                                 design's multi-wavefront kernel (one
                                 workgroup of 64 x rf_design.mw_waves threads
                                 per 64 stimuli) when it has one and no
                                 input stream or output capture is set;
                                 0 = one wavefront per 64 stimuli */
} rf_config;

typedef struct rf_model rf_model;

typedef struct rf_stats {
    uint64_t fused_launches;  /* fused / persistent kernel launches */
    double fused_ms;          /* their summed duration (HIP events on the model's stream) */
    uint64_t graph_launches;  /* graph replays */
    double graph_ms;          /* summed duration of graph replays (HIP events) */
    uint64_t bitplane_launches; /* bit-plane kernel launches */
    double bitplane_ms;       /* their summed duration (HIP events) */
} rf_stats;

/* Library / design ABI versions. */
uint32_t rf_abi_version(void);

/* Create a model: allocates the SoA state for n_lanes stimuli (zeroed: the
 * --x-initial 0 reset), the per-lane change/done/hash/cycle/stimulus arrays,
 * and builds the eval graph. Replaces RTLflow(size_t gpu_threads) +
 * RTLflow::initialize(Vtop__Syms*) (src/V3EmitC.cpp:5294, :5351-5366,
 * :5378-5497). cfg may be NULL for defaults. */
int rf_create(const rf_design* design, uint32_t n_lanes, const rf_config* cfg, rf_model** out);

/* Free everything (RTLflow::~RTLflow, src/V3EmitC.cpp:5367-5375). */
int rf_destroy(rf_model* m);

/* Port access for lanes [lane0, lane0+count). Host buffers hold, per lane,
 * ``words`` elements of the port's storage class (u8/u16/u32/u64; u32 words
 * for wide ports), lane-major. Replaces RTLflow::get(XDataLoc, idx) pointer
 * writes/reads (src/V3EmitC.cpp:5339-5350). Any port may be read. */
int rf_set_port(rf_model* m, uint32_t port, uint32_t lane0, uint32_t count, const void* src);
int rf_get_port(rf_model* m, uint32_t port, uint32_t lane0, uint32_t count, void* dst);

/* Memory (unpacked array) access for lanes [lane0, lane0+count), elements
 * [elem0, elem0+nelem): host buffer [count][nelem][words] of the class type. */
int rf_mem_write(rf_model* m, uint32_t mem, uint32_t lane0, uint32_t count, uint32_t elem0,
                 uint32_t nelem, const void* src);
int rf_mem_read(rf_model* m, uint32_t mem, uint32_t lane0, uint32_t count, uint32_t elem0,
                uint32_t nelem, void* dst);

/* Per-lane done mask (the reference's public ``done`` array, src/V3EmitC.cpp:5287-5292). */
int rf_set_done(rf_model* m, uint32_t lane0, uint32_t count, const uint8_t* done);

/* One eval() of every active lane with the settle loop (Vtop::eval ->
 * RTLflow::run, src/V3EmitC.cpp:3715, :5376, :5426-5496). */
int rf_eval(rf_model* m);

/* Clock-cycle protocol with on-device stimulus and hashing:
 * rf_begin sets each lane's hash to the FNV-1a-64 basis, its cycle counter to
 * 0 and its stimulus id to stim0 + lane; rf_run_cycles runs n cycles
 * (inputs for the lane's cycle with clk=0, eval, clk=1, eval, hash outputs).
 * mode RF_MODE_GRAPH replays the captured per-mtask graph once per cycle,
 * RF_MODE_FUSED runs the fused per-lane kernel, RF_MODE_BITPLANE runs the
 * bit-plane kernel (64 lanes per 64-bit word; RF_E_STATE if the design has
 * no bit-plane program). All three leave identical state and hashes.
 * rf_run_cycles queues its launches on the model's stream and returns; the
 * next call that waits for the device (rf_sync, rf_hashes, rf_get_port, ...)
 * reports their errors (RF_E_HIP, RF_E_CONVERGE) and adds their HIP-event
 * times to the statistics. */
int rf_begin(rf_model* m, uint64_t seed, uint32_t stim0);
int rf_run_cycles(rf_model* m, uint32_t n, uint32_t mode);
int rf_hashes(rf_model* m, uint32_t lane0, uint32_t count, uint64_t* out);

/* Batched form: run nstim stimuli with the given cycle budgets through the
 * model's lanes as a work queue in one persistent launch; each stimulus is
 * reset and initialised when it enters a lane. order may be NULL (longest
 * budget first). out_hashes[nstim] receives each stimulus's final hash. */
int rf_submit_queue(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim,
                    const uint32_t* order, uint64_t* out_hashes);

/* Largest batched workload rf_submit_queue_static takes. */
#define RF_QUEUE_STATIC_MAX 65536u

/* rf_submit_queue without allocating: no hipMalloc, hipFree or other
 * allocating call. No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it). The queue buffers (budgets, order, per-stimulus
 * hashes) live in a fixed device array of RF_QUEUE_STATIC_MAX stimuli
 * (16 B each, 1 MiB per GPU, held until the process ends); nstim above that
 * is refused with RF_E_ARG. Same arguments, results and errors as
 * rf_submit_queue. Every model on one GPU shares that GPU's array: calls
 * that overlap on one GPU (two host threads, two models) race. */
int rf_submit_queue_static(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim,
                           const uint32_t* order, uint64_t* out_hashes);

/* The batched workload in the bit-plane form (RF_E_STATE if the design has
 * no bit-plane program or memory images are set): the stimuli are dealt to
 * the lanes in queue order (order may be NULL: longest budget first), in
 * rounds of n_lanes; every round starts from the reset state, and each
 * 64-lane group runs as many cycles as its longest budget while every lane
 * stops hashing at its own. out_hashes[nstim] as rf_submit_queue. Lane state
 * afterwards is unspecified. No counterpart in the reference. */
int rf_submit_queue_bitplane(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim,
                             const uint32_t* order, uint64_t* out_hashes);

/* The batched workload in resumable steps. rf_queue_open sets up the queue
 * (order may be NULL: longest budget first) without running it;
 * rf_queue_step runs the next max_stimuli queue positions to completion in
 * ``mode`` (RF_MODE_FUSED: the persistent per-lane kernel; RF_MODE_BITPLANE:
 * bit-plane rounds) and returns the new position; rf_queue_hashes gives
 * every stimulus's hash so far (0 for those not yet run). Between steps no
 * lane holds a stimulus, so the queue's whole state is its budgets, order,
 * position and hashes: rf_queue_save writes them (rf_queue_save_bytes: the
 * size) and rf_queue_load restores them into any model of the same design,
 * whatever its lane count, to continue the batch -- the queue half of the
 * reference's --savable model save/restore (src/V3EmitC.cpp:2957-3039) for
 * long batches. Memory images and input streams are model settings: set
 * them again on the model that resumes. Hashes equal rf_submit_queue's. */
int rf_queue_open(rf_model* m, uint64_t seed, const uint32_t* budgets, uint32_t nstim, const uint32_t* order);
int rf_queue_step(rf_model* m, uint32_t max_stimuli, uint32_t mode, uint32_t* position);
int rf_queue_hashes(rf_model* m, uint64_t* out_hashes, uint32_t nstim);
int rf_queue_save_bytes(rf_model* m, size_t* bytes);
int rf_queue_save(rf_model* m, void* dst, size_t bytes);
int rf_queue_load(rf_model* m, const void* src, size_t bytes);

/* Per-stimulus memory images for the batched form: before a stimulus runs,
 * elements [0, nelem) of memory ``mem`` in its lane are loaded from
 * src[stimulus][nelem][words]. Pass src = NULL to clear. */
int rf_set_mem_images(rf_model* m, uint32_t mem, uint32_t nstim, uint32_t nelem, const void* src);

/* User stimulus and output capture for the clock-cycle protocol (every run
 * mode and the batched forms). The testbench side of the reference writes
 * each input through RTLflow::get into managed memory from the host
 * (src/V3EmitC.cpp:5339-5350) and reads outputs back the same way; here the
 * host hands over device buffers once and the kernels read and write them.
 *
 * Layout of both (rf_stream in rf_design.h): 64-bit words
 * [cycle - cycle0][word][stimulus - stim0], nstim stimuli per row.
 * rf_stream_words gives the words per stimulus-cycle: the input stream
 * holds every data input (clocks excluded) in declaration order, 64 bits per
 * word (port p's chunk j; masked to the port width when applied); the
 * capture holds every output likewise, taken at the point each cycle's hash
 * is folded. The design manifest lists the words ("streams").
 *
 * rf_set_input_stream replaces the SplitMix64 stimulus by ``words`` (device
 * memory of the model's GPU: hipMalloc or managed; NULL restores the
 * generator). A running lane that needs a stimulus-cycle outside the stream
 * is reported as RF_E_ARG by the next call that waits for the device (the
 * batched forms check their stimuli and budgets before launching).
 * rf_set_output_capture stores the outputs of the stimulus-cycles it covers
 * (partial windows are fine: the rest is not stored); NULL turns it off.
 * The buffers must stay valid while runs use them. */
int rf_stream_words(const rf_design* d, uint32_t* in_words, uint32_t* out_words);
int rf_set_input_stream(rf_model* m, const uint64_t* words, uint64_t stim0, uint32_t nstim,
                        uint32_t cycle0, uint32_t ncycles);
int rf_set_output_capture(rf_model* m, uint64_t* words, uint64_t stim0, uint32_t nstim,
                          uint32_t cycle0, uint32_t ncycles);

/* Raw state of one lane, every slot in class order C, S, I, Q then the
 * lane-major memory region (for lane-k state diffs against a CPU model,
 * the role of the reference's checkN asserter, src/V3EmitC.cpp:4380-4536).
 * *bytes receives the size; dst may be NULL to query it. */
int rf_lane_state(rf_model* m, uint32_t lane, void* dst, size_t* bytes);

/* Per-cycle state of one lane for a waveform dump: runs ``ncycles`` cycles
 * of the protocol in ``mode``, one launch (or graph replay) per cycle, and
 * after each copies lane ``lane``'s raw state (the rf_lane_state layout) on
 * the device; dst receives ncycles records of that size. The role of the
 * reference's VCD tracing of a model (src/V3EmitC.cpp:3927-3961,
 * include/verilated_vcd_c.cpp), for one lane of the batch. */
int rf_trace_lane(rf_model* m, uint32_t lane, uint32_t ncycles, uint32_t mode, void* dst, size_t bytes);

/* Checkpoint / resume of the protocol form: the whole simulation state of
 * a model (every SoA slot, the lane-major memories, the per-lane change,
 * done, hash, cycle-counter and stimulus-id arrays, the device flags and
 * the protocol seed) as one blob, so a long run can stop and continue in
 * another model of the same design and lane count -- the role of the
 * reference's --savable model save/restore (src/V3EmitC.cpp:2957-3039).
 * rf_checkpoint_bytes gives the blob size; dst / src may be host or device
 * memory. rf_checkpoint_save waits for queued work; rf_checkpoint_load
 * checks the header (design hash, lanes, layout) and queues its copies on
 * the model's stream. */
int rf_checkpoint_bytes(rf_model* m, size_t* bytes);
int rf_checkpoint_save(rf_model* m, void* dst, size_t bytes);
int rf_checkpoint_load(rf_model* m, const void* src, size_t bytes);

/* Time every mtask kernel of one eval in isolation (mean of ``reps``
 * back-to-back launches over all lanes, HIP events on the model's stream);
 * ms_out[nmtasks]. The measured costs train the learned cost model; no
 * counterpart in the reference (its --prof-threads instrumentation is
 * disabled in the GPU emitter, src/V3EmitC.cpp:1591-1613). */
int rf_profile_mtasks(rf_model* m, uint32_t reps, float* ms_out);

/* Wait for all queued work; device flags (RF_FLAG_*) of the last run. */
int rf_sync(rf_model* m);
int rf_flags(rf_model* m, uint32_t* out, uint32_t n);
int rf_get_stats(rf_model* m, rf_stats* out);
int rf_reset_stats(rf_model* m);

/* Lanes and padded lane stride of the model. */
uint32_t rf_lanes(const rf_model* m);
uint64_t rf_lane_stride(const rf_model* m);

const char* rf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
