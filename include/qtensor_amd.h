/*
 * qtensor_amd.h -- C ABI of the MI355X bucket-elimination engine.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses
 * owned by the caller; the library never frees caller memory.  Every entry
 * point returns QB_OK (0) or an error code; qb_last_error() gives the message
 * of the last failure on the calling thread.
 *
 * Reference interface each entry point replaces (paths in DaniloZZZ/QTensor):
 *   qb_step           -> BucketBackend.process_bucket   qtensor/ProcessingFrameworks.py:14-15
 *                        (NumpyBackend :29-31, CMKLExtendedBackend :45-147)
 *   qb_execute        -> the per-edge loop QAOASimulator.energy_expectation drives
 *                        (qtensor/QAOASimulator.py:44-66) around
 *                        bucket_elimination(process_bucket) (qtensor/Simulate.py:90-93),
 *                        one work item per (edge, slice) (FeynmanSimulator.py:49-58);
 *                        each item first builds, in FP64 from (gamma, beta), the gate
 *                        tensors that get_sliced_buckets materialises from data_dict
 *                        (ProcessingFrameworks.py:17-18,33-34)
 *   qb_exec_slots     -> (no reference counterpart: sizes the executor's scratch)
 *   qb_check_program  -> the structural checks of bucket elimination
 *                        (AssertionError at ProcessingFrameworks.py:74,137), on a
 *                        compiled program before it runs
 *   qb_wide_execute   -> bucket_elimination(process_bucket) over one large network
 *                        (qtensor/Simulate.py:90-93) or one slice of it
 *                        (qtensor/FeynmanSimulator.py:49-58), all steps on the device
 *   qb_plan_create / qb_plan_execute / qb_plan_destroy -> the same energy evaluation
 *                        with library-owned device buffers: compile once, re-evaluate
 *                        for new angles (QAOASimulator.energy_expectation,
 *                        QAOASimulator.py:44-66, over a fixed graph)
 *   qb_reduce_terms   -> sum of slice partials (FeynmanSimulator.py:83) and of edge
 *                        terms (QAOASimulator.py:60), in FP64
 *   qb_greedy_order   -> min-degree greedy elimination order (qtensor/utils.py:12-44,
 *                        used by OrderingOptimizer, qtensor/optimisation/Optimizer.py:46-88)
 *   qb_elimination_widths -> widths of a fixed order (the same elimination, utils.py:77-86)
 *   qb_rgreedy_order  -> randomized greedy orders, best of several samples
 *                        (RGreedyOptimizer, qtensor/optimisation/RGreedy.py:15-77)
 *   qb_place_intervals -> LDS / scratch offsets of a plan's buffers (the plan compiler's
 *                        own layout step; the reference allocates with numpy)
 *   qb_lightcone_classes / qb_compile_lightcone -> per-edge lightcone, circuit,
 *                        network, order, slices and plan (QAOASimulator.py:16-24,
 *                        Simulate.py:47-94, CircuitComposer.py:82-169), natively
 *   qb_init / qb_shutdown / qb_last_error -> backend construction and the Python
 *                        exceptions of the reference (AssertionError, ValueError)
 */
#ifndef QTENSOR_AMD_H
#define QTENSOR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QB_OK 0
#define QB_ERR_INVALID 1      /* bad argument or descriptor */
#define QB_ERR_HIP 2          /* HIP runtime error */
#define QB_ERR_UNSUPPORTED 3  /* shape outside what the kernels support */
#define QB_ERR_NOT_INIT 4     /* qb_init was not called */

#define QB_C64 0              /* complex64: interleaved float (re, im) */
#define QB_C128 1             /* complex128: interleaved double (re, im) */

#define QB_MAX_STEP_INPUTS 16 /* tensors per bucket step */
#define QB_MAX_RANK 31        /* size-2 axes in a step's index union */

/* Words per work item in qb_exec_desc.items. */
#define QB_ITEM_WORDS 4
/* item[0] = offset of the item's plan in the program buffer (int32 words)
 * item[1] = offset of the item's angle vector in params (doubles)
 * item[2] = slice assignment: bit k = value of the plan's k-th sliced variable
 * item[3] = output slot: the item's value goes to item_out[2*slot .. 2*slot+1]
 *           (slots are a permutation of 0..n_items-1; items may be listed in any
 *           order, e.g. longest first) */

/* Library version (major*10000 + minor*100 + patch). */
int qb_version(void);

/* Select device `device` and record its properties.  Idempotent. */
int qb_init(int device);

/* Release library state (device buffers it owns).  Idempotent. */
int qb_shutdown(void);

/* Message of the last failed call on this thread ("" if none). */
const char* qb_last_error(void);

/*
 * One bucket step: multiply the step's inputs over the union U of their
 * size-2 indices (ascending variable id, C order) and, if `sum`, sum over the
 * first union axis (the bucket variable).
 *
 *   out[o] = sum_{x in {0,1} if sum} prod_j in_j[ pext(x*2^(rank_u-1) + o, masks[j]) ]
 *
 * where pext gathers the bits of the union offset selected by masks[j]:
 * bit (rank_u-1-a) of masks[j] is set iff input j carries union axis a.
 * Size-1 axes carry no data and appear in no mask.  `out` holds
 * 2^(rank_u - sum) elements and must not alias an input.
 */
typedef struct qb_step_desc {
    int32_t n_inputs;
    int32_t rank_u;
    int32_t sum;
    int32_t dtype;                         /* QB_C64 or QB_C128 */
    uint32_t masks[QB_MAX_STEP_INPUTS];
} qb_step_desc;

int qb_step(const qb_step_desc* desc, const void* const* inputs, void* out, void* stream);

/*
 * Batched executor: every work item contracts one compiled lightcone plan
 * (one edge, or one slice of one edge) to a scalar.  A persistent grid runs
 * item b in block b first and pulls later items from `queue`; each item builds its
 * gate tensors in LDS from its FP64 angles and runs the plan's steps with
 * intermediates in LDS or, above the plan's LDS threshold, in its slot of
 * `scratch`.  Writes one complex128 (re, im) per item to item_out.
 */
typedef struct qb_exec_desc {
    const int32_t* program;      /* device: concatenated plans */
    const int32_t* items;        /* device: n_items * QB_ITEM_WORDS */
    int32_t n_items;
    int32_t item_threads;        /* 64, 128, 256, 512 or 1024 */
    int32_t lds_elems;           /* LDS elements reserved per item (>= every plan's need) */
    int32_t dtype;               /* QB_C64 or QB_C128 */
    const double* params;        /* device: angle vectors (FP64); readable for 8 doubles
                                    past every item's offset (pad the buffer) */
    const void* const_pool;      /* device: constant gate data, dtype elements */
    void* scratch;               /* device: n_slots * scratch_elems elements */
    int64_t scratch_elems;       /* elements per slot (>= every plan's global need) */
    int32_t n_slots;             /* slots available in scratch */
    int32_t program_words;       /* words of program, at most 2^30: op records and
                                    lane tables are read through a buffer of this
                                    size (reads past it return 0); 0 = unbounded */
    double* item_out;            /* device: 2 * n_items doubles */
    uint32_t* queue;             /* device: one uint32 (reset by the call when the
                                    groups outnumber the resident blocks) */
    int32_t group;               /* items per block, run in lockstep (0 or 1: one;
                                    2: complex64 with item_threads 64 ... 512).
                                    Items come in groups of `group` consecutive
                                    items sharing one plan offset; n_items is a
                                    multiple of group.  A group keeps its items'
                                    LDS data (and scratch) interleaved, so each
                                    block reserves group * lds_elems elements and
                                    group scratch slots. */
} qb_exec_desc;

/* Number of concurrently resident work items (= scratch slots needed) for a
 * launch shape with `group` items per block (0 or 1: one); qb_execute uses at
 * most this many slots. */
int qb_exec_slots(int32_t item_threads, int32_t lds_elems, int32_t dtype, int32_t group, int32_t* n_slots);

/* qb_execute checks the descriptor's sizes, but it reads the items and the
 * program through raw device pointers and does NOT validate their contents:
 * a plan offset outside the program, a plan needing more LDS or scratch than
 * the descriptor reserves, or an output slot outside item_out makes the
 * kernel read or write out of bounds.  Run qb_check_program on host copies of
 * the same program and items first (qb_plan_create always does). */
int qb_execute(const qb_exec_desc* desc, void* stream);

/*
 * qb_execute without caller workspace.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * descriptor is qb_execute's with scratch = NULL, n_slots = 0 and queue =
 * NULL: the work-queue ticket and the items' global scratch slots live in a
 * library-owned static device array of QB_NOALLOC_SCRATCH_BYTES per device
 * (256 bytes for the ticket, the rest for scratch), resolved by qb_init.  The
 * call allocates nothing and rejects a launch whose resident items x
 * scratch_elems elements do not fit (QB_ERR_UNSUPPORTED).  Every call on a
 * device shares that array, so calls that overlap on one device race; the
 * array keeps the last call's contents (the ticket is reset by a fill kernel
 * whenever a launch takes tickets).  item_out, the program, items and angles
 * stay the caller's, unchecked as for qb_execute.
 */
#define QB_NOALLOC_SCRATCH_BYTES (64u << 20)
int qb_execute_noalloc(const qb_exec_desc* desc, void* stream);

/*
 * Host check of a batched program and its items (no device work; usable
 * without qb_init): every item's plan offset, angle offset and output slot
 * (slots must be a permutation of 0..n_items-1), with `group` items per
 * block (0 or 1: one) that every group's items share a plan offset and that
 * n_items is a multiple of group, and for every plan its
 * header, gate table (device-built kinds only, inside the LDS data, angle
 * indices inside the item's angle vector), every op record and lane table
 * (inside the program, every LDS element below the plan's LDS need, which
 * must not exceed lds_elems, every scratch location below its scratch need,
 * which must not exceed scratch_elems, every mask consistent with its run
 * word and summed-variable offset) and its result location.
 * Replaces the structural assertions the reference makes while eliminating
 * (ProcessingFrameworks.py:74,137).
 */
int qb_check_program(const int32_t* program, int64_t program_words, const int32_t* items, int32_t n_items,
                     int32_t n_params, int32_t item_threads, int32_t lds_elems, int64_t scratch_elems,
                     int32_t group);

/*
 * Wide plan: one compiled lightcone plan whose tensors (gates included) all
 * live in `arena` (global memory), for networks too large for the executor's
 * LDS tiers.  Builds the gate tensors from `params` for the slice assignment
 * `slice_bits`, launches one grid-wide kernel per op in order on `stream` (a
 * bucket step, a fused chain of steps, or a chain op: a contraction and the
 * contraction reading its output, whose intermediate is never stored; the
 * matrix cores form its inner sums when it has the tile shape they need),
 * and writes the result as complex128: the scalar to out[0..1], or, for a
 * plan with b open (batch) variables (plan word 11; the amplitudes of
 * `QtreeSimulator.simulate_batch(batch_vars=b)`, qtensor/Simulate.py:66-94),
 * element e of the 2^b result to out[2 e out_stride .. + 1] (first open
 * variable = most significant bit of e).  The host copy of the plan is parsed
 * and every tensor range is checked against `arena_elems` before any launch.
 * With n_slices = S > 1, one call evaluates the S slice assignments
 * slice_bits .. slice_bits + S - 1 in lockstep: every launch covers all S
 * (the grid's z dimension), slice z uses arena elements
 * [z arena_elems, (z + 1) arena_elems) (the arena then holds S * arena_elems
 * elements) and writes result element e to out[2 (e out_stride + z) .. + 1]
 * (a plan with open variables needs out_stride >= S).  One call per S slices
 * instead of S calls: S times fewer launches, each S times wider.
 */
typedef struct qb_wide_desc {
    const int32_t* plan_host;    /* host: plan words */
    int64_t plan_words;
    const int32_t* plan_dev;     /* device: the same words */
    const double* params;        /* device: angle vector (FP64) */
    void* arena;                 /* device: arena_elems elements of dtype */
    int64_t arena_elems;
    uint32_t slice_bits;
    int32_t dtype;
    double* out;                 /* device: 2 doubles per result element */
    int64_t out_stride;          /* complex128 values between result elements (0 or 1: adjacent) */
    int32_t n_slices;            /* slices per call (0 or 1: one; at most 1024) */
} qb_wide_desc;

int qb_wide_execute(const qb_wide_desc* desc, void* stream);

/*
 * FP64 reductions: segment s covers items [seg[s], seg[s+1]); term_out[s] is
 * the (Neumaier-compensated) complex sum of its items' values, and total[0..1]
 * the complex sum of all terms.  n_segments <= 2^20.  seg = NULL means one
 * item per segment (item s is term s): one launch instead of two.
 */
int qb_reduce_terms(const double* item_out, const int32_t* seg, int32_t n_segments,
                    double* term_out, double* total, void* stream);

/*
 * Plan handles (owned device state).  qb_plan_create uploads a batched program
 * (concatenated compiled plans followed by at least 320 zero words), its work
 * items (QB_ITEM_WORDS each, as for qb_execute) and the term segments
 * (n_terms + 1 offsets into the items), and allocates item outputs, scratch
 * slots and the work queue on the current device.  Items and segments are
 * checked on the host.  qb_plan_execute runs the executor and the FP64
 * reductions on `stream` for one device angle buffer `params` (readable for
 * n_params + 8 doubles): term_out gets 2 * n_terms doubles (complex terms),
 * total 2 doubles.  Calls on one handle must not overlap (they share its
 * buffers).  qb_plan_destroy frees everything (null is allowed).
 */
typedef struct qb_plan qb_plan;

typedef struct qb_plan_desc {
    const int32_t* program;      /* host: program words */
    int64_t program_words;
    const int32_t* items;        /* host: n_items * QB_ITEM_WORDS */
    int32_t n_items;
    int32_t n_terms;
    const int32_t* seg;          /* host: n_terms + 1 item offsets */
    int32_t n_params;            /* doubles of one angle buffer (without padding) */
    int32_t item_threads;        /* 64, 128, 256, 512 or 1024 */
    int32_t lds_elems;           /* LDS elements per item (>= every plan's need) */
    int32_t dtype;               /* QB_C64 or QB_C128 */
    int64_t scratch_elems;       /* global scratch elements per resident item */
    int32_t group;               /* items per block, as in qb_exec_desc; the
                                    segments may end before n_items (padding items
                                    that complete a group and belong to no term) */
} qb_plan_desc;

int qb_plan_create(const qb_plan_desc* desc, qb_plan** plan);
int qb_plan_execute(qb_plan* plan, const double* params, double* term_out, double* total, void* stream);
int qb_plan_destroy(qb_plan* plan);

/*
 * Host-only planning (no device work; usable without qb_init).
 * A graph of n_nodes nodes 0..n_nodes-1 in the caller's node order is given in
 * CSR form: the neighbours of node v are adj_idx[adj_ptr[v] .. adj_ptr[v+1]).
 * Self loops and repeated neighbours are ignored; edges are taken as undirected.
 *
 * qb_greedy_order: repeatedly eliminate the node with the fewest neighbours
 * (the first such node in node order) and join its neighbours into a clique;
 * order[k] is the k-th eliminated node and widths[k] its neighbour count then.
 */
int qb_greedy_order(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, int32_t* order,
                    int32_t* widths);

/* Widths of eliminating order[0..n_order) in that order (same elimination rule). */
int qb_elimination_widths(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, int32_t n_order,
                          const int32_t* order, int32_t* widths);

/*
 * qb_rgreedy_order (replaces RGreedyOptimizer._get_ordering_ints,
 * qtensor/optimisation/RGreedy.py:36-77): samples 1..repeats eliminate, at
 * every step, a remaining node drawn with probability proportional to
 * exp(-(neighbours - fewest neighbours) / temp) (temp > 0; a splitmix64 stream
 * seeded by `seed`); sample 0 is the qb_greedy_order elimination.  order and
 * widths receive the sample with the smallest (largest width, sum of
 * 2^width), the earliest on ties.  Deterministic for a given seed.
 */
int qb_rgreedy_order(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, double temp,
                     uint64_t seed, int32_t repeats, int32_t* order, int32_t* widths);

/*
 * Interval allocation for the plan compiler (host only; no counterpart in the
 * reference, which allocates every intermediate with numpy): buffer i lives
 * over ops [starts[i], ends[i]] and needs sizes[i] elements.  In `order`,
 * each buffer gets the lowest offset >= base free over its whole interval
 * among the buffers already placed; *peak = the highest end (>= base).
 */
int qb_place_intervals(int32_t m, const int64_t* sizes, const int64_t* starts, const int64_t* ends,
                       const int64_t* order, int64_t base, int64_t* offsets, int64_t* peak);

/*
 * Native lightcone planner (host only; usable without qb_init).
 *
 * qb_lightcone_classes replaces the per-edge lightcone extraction of
 * QtreeQAOAComposer.energy_expectation_lightcone (qtensor/CircuitComposer.py:82-96,
 * qtensor/utils.py:176-189) for a whole graph: every edge's radius-p ball is
 * relabelled canonically (colour refinement seeded with endpoint and distance,
 * then individualisation), and edges whose relabelled lightcones are equal share
 * a class.  The graph (nodes 0..n_nodes-1) is CSR as for qb_greedy_order; edges
 * holds n_edges (u, v) pairs.  edge_class[e] receives the class of edge e;
 * class_words receives, per class in order of first appearance,
 * [n, n_edges_c, i, j, a0, b0, a1, b1, ...] (canonical labels, a < b, sorted;
 * (i, j) the edge's own endpoints).  If capacity is too small the call fails with
 * QB_ERR_UNSUPPORTED and *used tells the size needed.
 */
int qb_lightcone_classes(int32_t n_nodes, const int32_t* adj_ptr, const int32_t* adj_idx, int32_t n_edges,
                         const int32_t* edges, int32_t p, int32_t* edge_class, int32_t* n_classes,
                         int32_t* class_words, int64_t capacity, int64_t* used);

/*
 * qb_compile_lightcone replaces, for one lightcone class, the chain that
 * QAOASimulator._get_edge_energy and QtreeSimulator.simulate_batch run per edge
 * (qtensor/QAOASimulator.py:16-24, qtensor/Simulate.py:47-94): the QAOA circuit
 * U Z_i Z_j U^dagger with cost gates as cX.ZPhase.cX (composer 0,
 * QtreeQAOAComposer), ZZ (1, ZZQtreeQAOAComposer) or ZZ with commutation-based
 * cancellation (2, DefaultQAOAComposer; qtensor/__init__.py:55-63), the bucket
 * network, the elimination order (0: min-degree greedy, qtensor/utils.py:12-44;
 * 1: best of rg_repeats randomized-greedy samples and the greedy order,
 * RGreedy.py:15-77), slices (0: none; 1: GreedyParvars, Optimizer.py:90-159;
 * 2: tree trimming, Optimizer.py:178-228) until the width is below
 * max_tw - tw_bias, and the executor plan of that network (qb_execute's program
 * format) for work items of `lanes` threads, intermediates of at most
 * lds_tensor_limit elements in LDS within lds_capacity elements (-1: no limit).
 * `lightcone` is one class record of qb_lightcone_classes.  Errors: an LDS
 * capacity that the plan exceeds (QB_ERR_INVALID, "LDS capacity"), a width the
 * slicer cannot reach (QB_ERR_UNSUPPORTED, "Estimated OOM", as the reference's
 * Optimizer.py:124).
 */
typedef struct qb_lightcone_plan_opts {
    int32_t composer;          /* 0, 1 or 2 (above) */
    int32_t ordering;          /* 0 greedy, 1 randomized greedy */
    int32_t slicer;            /* 0 none, 1 GreedyParvars, 2 tree trimming */
    int32_t max_tw;
    int32_t tw_bias;
    int32_t lanes;             /* 64, 128, 256, 512 or 1024 */
    int64_t lds_tensor_limit;
    int64_t lds_capacity;      /* -1: unbounded */
    double rg_temp;            /* randomized greedy: temperature, samples, seed */
    int32_t rg_repeats;
    uint64_t rg_seed;
} qb_lightcone_plan_opts;

typedef struct qb_lightcone_plan_info {
    int32_t n_steps, n_ops, n_gates, treewidth, max_rank, n_sliced;
    int64_t lds_need, glb_need;
    int64_t bytes_step_model, flops;   /* complex64 step model of the bucket steps */
    int32_t n_barriers, n_gather, n_reduce, steps_words;
} qb_lightcone_plan_info;

int qb_compile_lightcone(const int32_t* lightcone, int32_t p, const qb_lightcone_plan_opts* opts, int32_t* words,
                         int64_t capacity, int64_t* used, qb_lightcone_plan_info* info);

#ifdef __cplusplus
}
#endif

#endif /* QTENSOR_AMD_H */
