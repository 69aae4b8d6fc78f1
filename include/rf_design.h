/* rf_design.h -- ABI between a generated design library and the runtime.
 *
 * A design compiled by rtlflow_amd (the V3EmitC replacement) is a shared
 * library that exports one function, ``rf_design_get()``, returning a static
 * ``rf_design`` table: the SoA layout sizes, the port table (the reference's
 * ``RF_IN*``/``RF_OUT*`` DataLoc port descriptors, include/rf_heavy.h:19-51),
 * the mtask DAG (the reference's Taskflow cudaFlow built in
 * ``RTLflow::initialize``, src/V3EmitC.cpp:5378-5424) and the kernels.
 *
 * Every kernel takes one ``rf_args`` by value. It replaces the reference's
 * kernel ABI ``(void* symtab, CData*, SData*, IData*, QData*, IData* change,
 * bool* done)`` (src/V3EmitC.cpp:1648-1649) with explicit lane stride and the
 * per-lane protocol state (cycle counter, stimulus id, hash).
 *
 * Plain C: no HIP or torch types.
 */
#ifndef RF_DESIGN_H
#define RF_DESIGN_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_DESIGN_ABI 8

/* storage classes (include/rf_heavy.h:11-17: CData/SData/IData/QData) */
enum { RF_CLS_C = 0, RF_CLS_S = 1, RF_CLS_I = 2, RF_CLS_Q = 3 };
/* port directions */
enum { RF_PORT_IN = 0, RF_PORT_OUT = 1, RF_PORT_CLOCK = 2 };
/* memory layouts */
enum { RF_MEM_ELEMENT_MAJOR = 0, RF_MEM_LANE_MAJOR = 1 };
/* eval-end phases (rf_args.phase) */
enum { RF_PHASE_EVAL = 0, RF_PHASE_RISE = 1, RF_PHASE_HASH = 2 };
/* device flag words (rf_args.flags[]) */
enum {
    RF_FLAG_CHANGE = 0,       /* OR over lanes of this eval's change request */
    RF_FLAG_UNCONVERGED = 1,  /* a lane hit the settle limit */
    RF_FLAG_QUEUE = 2,        /* batched form: next stimulus to hand out */
    RF_FLAG_ACTIVE = 3,       /* batched form: lanes still running */
    RF_FLAG_STREAM = 4,       /* a running lane read outside the input stream */
    RF_FLAG_WORDS = 8
};

typedef struct rf_port {
    const char* name;
    uint32_t dir;     /* RF_PORT_* */
    uint32_t width;   /* bits */
    uint32_t cls;     /* RF_CLS_*; wide ports are RF_CLS_I with words > 1 */
    uint32_t words;   /* storage slots */
    uint64_t loc;     /* first slot in its class array */
    uint32_t stim_index; /* data inputs: index p of the stimulus stream; else ~0u */
    uint32_t hashed;  /* outputs: folded into the per-cycle hash */
} rf_port;

typedef struct rf_mem {
    const char* name;
    uint32_t width, cls, words, depth;
    uint32_t layout;  /* RF_MEM_* */
    uint64_t loc;     /* element-major: first slot; lane-major: byte offset factor */
    uint64_t row_bytes; /* lane-major: bytes per lane */
} rf_mem;

typedef struct rf_mtask {
    const void* kernel;     /* __global__ void k(rf_args) host stub */
    const char* name;
    uint32_t npreds;
    const uint32_t* preds;  /* indices of predecessor mtasks */
    uint32_t cost;          /* estimated instruction count */
} rf_mtask;

/* A device buffer of 64-bit words laid out [cycle][word][stimulus]: row
 * ``c - cycle0`` holds, for every port word ``w`` in the design's port order,
 * ``nstim`` consecutive stimuli from ``stim0``. As the input stream it
 * replaces the SplitMix64 stimulus of the clock-cycle protocol (``words`` =
 * the design's input words: data inputs in declaration order, 64 bits per
 * word, masked to the port width); as the output capture it receives every
 * output word at the point the cycle's hash is taken. Null data: unused. */
typedef struct rf_stream {
    uint64_t* data;
    uint64_t stim0;
    uint32_t nstim;
    uint32_t cycle0;
    uint32_t ncycles;
    uint32_t words;
} rf_stream;

typedef struct rf_args {
    uint8_t* c;
    uint16_t* s;
    uint32_t* i;
    uint64_t* q;
    uint8_t* mem;        /* lane-major memory region */
    uint32_t* change;    /* [npad] per-lane change request */
    uint8_t* done;       /* [npad] per-lane mask (user-settable) */
    uint64_t* hash;      /* [npad] per-lane FNV-1a-64 state */
    uint32_t* cyc;       /* [npad] per-lane cycle counter */
    uint32_t* stim;      /* [npad] per-lane stimulus id */
    uint32_t* flags;     /* [RF_FLAG_WORDS] */
    const uint32_t* budget; /* batched: [nstim] cycles per stimulus */
    const uint32_t* order;  /* batched: [nstim] queue order (stimulus ids) */
    uint64_t* out_hash;  /* batched: [nstim] final hash per stimulus */
    const uint8_t* images; /* batched: per-stimulus memory images (may be null) */
    uint64_t npad;       /* lane stride of every slot */
    uint64_t seed;
    uint32_t nlanes;
    uint32_t nstim;
    uint32_t ncycles;    /* fused: cycles per launch */
    uint32_t phase;      /* eval-end: RF_PHASE_* */
    uint32_t converge_limit;
    uint32_t pad0;
    rf_stream in;        /* user stimulus (data null: SplitMix64) */
    rf_stream out;       /* output capture (data null: none) */
} rf_args;

typedef struct rf_design {
    uint32_t abi;              /* RF_DESIGN_ABI */
    uint32_t block;            /* threads per block the kernels assume */
    const char* name;
    uint64_t hash;             /* hash of the generated source */
    uint64_t slots[4];         /* per-lane slots in the C/S/I/Q arrays */
    uint64_t mem_row_bytes;    /* per-lane bytes of the lane-major region */
    uint32_t nports;
    const rf_port* ports;
    uint32_t nmems;
    const rf_mem* mems;
    uint32_t nmtasks;
    const rf_mtask* mtasks;    /* one eval's DAG, topologically ordered */
    const void* k_init;        /* initial blocks + clock latch + settle */
    const void* k_eval_end;    /* last-assign + change request (+ phase work) */
    const void* k_stim;        /* protocol: cycle inputs, clk = 0 */
    const void* k_fused;       /* protocol: ncycles full cycles per lane */
    const void* k_persistent;  /* batched: work queue, refill on finish */
    const void* k_reset;       /* zero one lane's state (for refill / reset) */
    const void* k_settle;      /* graph form: lanes still changing after the
                                  captured passes finish the settle loop
                                  (NULL when nchange == 0); rf_args.ncycles =
                                  passes already run */
    /* k_fused / k_persistent with the input-stream and output-capture code:
     * the runtime launches these when rf_args.in or rf_args.out is set (the
     * plain kernels keep that code out of their cycle loops) */
    const void* k_fused_io;
    const void* k_persistent_io;
    uint32_t nchange;          /* change-detected variables (0: one pass settles) */
    uint32_t clk_port;         /* index into ports, or ~0u */
    /* algorithmic bytes per stimulus-cycle (layout manifest) */
    uint64_t bytes_seq, bytes_mem, bytes_in, bytes_out;
    uint64_t ops_per_cycle;    /* estimated instructions per stimulus-cycle */
    /* bit-plane form: the design lowered to single-bit logic for 64 lanes
     * per 64-bit word (NULL when the design is outside that subset) */
    const uint32_t* bitplane;
    uint64_t bitplane_words;
    /* multi-wavefront fused kernel: one workgroup of 64 x mw_waves threads
     * per 64 stimuli, each wavefront running its share of every level of the
     * mtask DAG (NULL / 0 when the design has none); the runtime launches it
     * for rf_run_cycles in RF_MODE_FUSED when rf_config.fused_multiwave is
     * set and no input stream or output capture is */
    const void* k_fused_mw;
    uint32_t mw_waves;
} rf_design;

typedef const rf_design* (*rf_design_get_fn)(void);

#ifdef __cplusplus
}
#endif
#endif
