/*
 * lpa_amd.h -- C ABI of the MI355X (gfx950) label-propagation engine,
 * liblpa_hip.so.  Plain pointers and sizes; no C++ or torch types.
 *
 * The reference's boundary is in-process C++ (the baselines/<method> propagators
 * launching CUDA kernels).  Each entry point below names the reference
 * interface it replaces:
 *
 *   lpa_engine_create/destroy   init_gmem / free_gmem
 *                               (baselines/nova/nova_propogator.cuh:313-345, 381-394)
 *   lpa_graph_upload            transfer_data, the H2D copy of the CGR words,
 *                               offsets, vertex order and small CSR
 *                               (nova_propogator.cuh:347-374)
 *   lpa_graph_degrees           the CSR degrees the binning reads
 *                               (utils.h:44-85 reads offsets[v+1]-offsets[v])
 *   lpa_schedule_degree         tiny_node_assignment + binary_search binning
 *                               (nova_propogator.cuh:210-311; utils.h:44-85)
 *   lpa_schedule_bins           (no counterpart) explicit per-vertex bins, used
 *                               by the learned schedule
 *   lpa_lp_config               compile-time CMS / Decoder parameters of the
 *                               reference (cms.cuh:7; nova/kernel.cuh:28)
 *   lpa_labels_init             initialize_labels (nova/kernel.cuh:57-64)
 *   lpa_labels_get              the D2H copy of labels (nova_propogator.cuh:123)
 *   lpa_nova_iterate            NovaPlusPropagator::iterate: propagate_large_vertex,
 *                               propagate_mid_vertex, l_small_*, l_tiny_*
 *                               (baselines/novaplus/novaplus_propogator.cuh:131-199;
 *                                baselines/novaplus/kernel.cuh:14-191)
 *   lpa_cglp_decode             decompress_large_vertex / decompress_mid_vertex
 *                               (baselines/cglp/kernel.cuh:10-39)
 *   lpa_cglp_iterate            CGLPPropagator::iterate: decode, then the
 *                               gather/sort/scan/compute_max_labels LP stage
 *                               (baselines/cglp/cglp_propagator.cuh:205-335)
 *   lpa_csr_iterate             the uncompressed-CSR comparator of baselines/g-sort
 *                               (InCoreLIDPP::iterate, incore_li.cuh:132-186: LP
 *                               over the plain CSR), here with the hash kernels
 *   lpa_cglp_csr_get            (no counterpart) copy of the decoded CSR, for checks
 *   lpa_time_iterations         the per-iteration Timer around iterate
 *                               (nova_propogator.cuh:114-125), with HIP events
 *
 * Semantics ("LP-SC"): synchronous label propagation with double-buffered
 * labels; a vertex takes the most frequent label among its neighbours, ties
 * to the smallest label; no (labelled) neighbour -> keeps its label.  With
 * CMS enabled, vertices of degree > cms_thresh rank labels by a count-min
 * estimate (depth d, width w, multiply-shift hashes from splitmix64(seed)).
 * Semi-supervised mode: 0xFFFFFFFF = unlabelled (ignored), initially
 * labelled entries stay clamped.  Labels must be < 2^31 (or 0xFFFFFFFF).
 *
 * Conventions: every function returns LPA_OK (0) or a negative status and
 * leaves a message for lpa_last_error() (thread-local).  Device memory is
 * owned by the engine handle.  One host thread per engine; all work is
 * queued on the engine's own stream (default priority).
 */
#ifndef LPA_AMD_H
#define LPA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LPA_OK = 0,
    LPA_ERR_ARG = -1,       /* bad argument */
    LPA_ERR_HIP = -2,       /* HIP runtime error */
    LPA_ERR_FORMAT = -3,    /* malformed CGR stream or offsets */
    LPA_ERR_STATE = -4,     /* call out of order (no graph, no labels, ...) */
    LPA_ERR_RANGE = -5      /* size or value out of the supported range */
};

typedef struct lpa_engine lpa_engine;

/* kernel families a vertex can be assigned to */
enum {
    LPA_BIN_SKIP = 0,       /* degree 0: label never changes */
    LPA_BIN_LANE = 1,       /* one lane per vertex, register sort (deg <= 32) */
    LPA_BIN_WAVE = 2,       /* one wave per vertex, LDS hash */
    LPA_BIN_BLOCK = 3,      /* one workgroup per vertex, LDS hash */
    LPA_BIN_GLOBAL = 4,     /* heavy exact: one 1024-lane workgroup per vertex (LDS table), or
                               several per hub (chunks, HBM table) */
    LPA_BIN_CMS = 5,        /* heavy count-min sketch (forced by degree): one workgroup per vertex
                               (LDS sketches), or several per hub */
    LPA_NBIN_KINDS = 6
};

typedef struct {
    uint32_t lane_max;      /* deg <= lane_max -> LANE (<= 64) */
    uint32_t wave_max;      /* deg <= wave_max -> WAVE (0: 512) */
    uint32_t block_max;     /* deg <= block_max -> BLOCK, else the heavy-vertex path (0: LDS limit) */
    uint32_t relayout;      /* records are always reordered so each bin is contiguous; keep 1 */
    uint32_t quarter_max;   /* WAVE vertices up to this degree use 16-lane groups (0: 128) */
    uint32_t bin_order;     /* vertices inside a bin: 1 by vertex id (neighbouring ids run together),
                               2 by degree, descending; 0: the engine's default (2) */
    uint32_t xcd_tiles;     /* 1: the workgroups one XCD runs take consecutive tiles of a bin (each XCD
                               has its own L2), 2: tiles in workgroup order; 0: the engine's default (2) */
} lpa_sched_params;

const char *lpa_last_error(void);
int lpa_device_count(int *count);

int lpa_engine_create(int device, lpa_engine **out);
int lpa_engine_destroy(lpa_engine *e);
/* info[0..3] = compute units, LDS bytes per workgroup, largest degree of the
 * WAVE and of the BLOCK kernels' LDS tables. */
int lpa_engine_info(lpa_engine *e, uint64_t *info);

/* Graph in the reference's File layout: words are the big-endian 32-bit
 * words of File::graph (nwords of them), offsets the n_c+1 START bit offsets
 * (File::offset, 64-bit), vertices the n ids (stream order, then small),
 * small_offsets/small_nbrs the CSR of the n-n_c small vertices (may be
 * NULL when n_c == n).  The stream is validated on the device: every record
 * must decode within its offsets to neighbour ids < n. */
int lpa_graph_upload(lpa_engine *e, uint64_t n, uint64_t n_c,
                     const uint32_t *words, uint64_t nwords, const uint64_t *offsets,
                     const uint32_t *vertices, const uint64_t *small_offsets,
                     const uint32_t *small_nbrs);
/* Encodes a CSR (host arrays: row[n+1], col[row[n]], every row strictly
 * increasing) into the CGR stream on the device and makes it the engine's
 * graph, as lpa_graph_upload would with the same file set: vertex order
 * deg >= 128 by id, then small_thresh <= deg < 128 by id, deg <
 * small_thresh as the small CSR.  Replaces CgrCompressor::encode_node /
 * write_cgr (src/compressor.cpp:157-515) run on the host; the bitstream is
 * the reference compressor's, bit for bit.  Rejects unsorted rows,
 * duplicate or out-of-range ids (LPA_ERR_FORMAT). */
int lpa_graph_encode_csr(lpa_engine *e, uint64_t n, const uint64_t *row, const uint32_t *col,
                         uint32_t small_thresh);
/* BFS relabelling of a CSR on the device: host arrays row[n+1], col[row[n]]
 * (ids < n) in; new_row[n+1], new_col[row[n]] and perm[n] (perm[v] = the new
 * id of v) out.  New ids follow a breadth-first search from `root` taking
 * neighbours in ascending order, then the unreached vertices by id; every
 * relabelled row is sorted.  Replaces the host's lpah_csr_bfs_relabel
 * (include/lpa_host.h), with the same result bit for bit; the reference
 * takes its vertex order from the input file (src/file.cpp:115-195) and has
 * no relabelling of its own. */
int lpa_csr_bfs_relabel(lpa_engine *e, uint64_t n, const uint64_t *row, const uint32_t *col, uint32_t root,
                        uint64_t *new_row, uint32_t *new_col, uint32_t *perm);
/* Allocation-free variant of lpa_graph_encode_csr.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * All arrays are DEVICE pointers the caller owns: row[n+1], col[row[n]]
 * (every row strictly increasing), vid[n_c] the vertex of each compressed
 * position (the stream order, e.g. deg >= 128 by id, then the rest by id);
 * outputs offsets[n_c+1] (start bit of each record, offsets[n_c] = total)
 * and words[words_cap] (64-bit words, bit 0 = MSB of word 0; the first
 * ceil(total/64) are written).  *nbits (host) receives the total.  The same
 * kernels as lpa_graph_encode_csr write the same bits; nothing is allocated:
 * the scan's block sums and the error flag live in a 1 MiB device array of
 * the library (one per GPU, kept until the process ends), which limits n_c
 * to 536,866,816 records (LPA_ERR_RANGE beyond).  words_cap too small ->
 * LPA_ERR_RANGE with words untouched; a bad row -> LPA_ERR_FORMAT.  Runs on
 * the engine's stream and returns when done; calls that overlap on one GPU
 * (several threads or engines) race on the shared array. */
int lpa_encode_csr_device(lpa_engine *e, uint64_t n, uint64_t n_c, const uint64_t *row, const uint32_t *col,
                          const uint32_t *vid, uint64_t *offsets, uint64_t *words, uint64_t words_cap,
                          uint64_t *nbits);
/* The engine's compressed part as stored: *n_c, *nbits, and when not NULL
 * words[ceil(nbits/64)] (64-bit words, bit 0 = MSB of word 0), offsets[n_c+1]
 * (start bit of each record) and vertices[n] (stream order, then small) --
 * what lpah_files_write needs to write the reference's file set. */
int lpa_graph_stream(lpa_engine *e, uint64_t *n_c, uint64_t *nbits, uint64_t *words, uint64_t *offsets,
                     uint32_t *vertices);
/* deg[v] for every vertex id v (n entries), computed by decoding. */
int lpa_graph_degrees(lpa_engine *e, uint32_t *deg);
/* Per-vertex work features for the cost model, by vertex id, 8 floats each:
 * degree, #intervals, #residuals, encoded bits, #interval segments,
 * #residual segments, mean run length, log2(1+degree). */
int lpa_graph_features(lpa_engine *e, float *feat);

/* LP parameters; changing K or the CMS parameters drops the schedule, so
 * call this before lpa_schedule_*. */
int lpa_lp_config(lpa_engine *e, uint32_t K, int semi, uint32_t cms_thresh,
                  uint32_t cms_depth, uint32_t cms_width, uint64_t cms_seed);
/* Degree-threshold schedule; CMS vertices (deg > cms_thresh) always go to
 * LPA_BIN_CMS whatever the schedule says. */
int lpa_schedule_degree(lpa_engine *e, const lpa_sched_params *p);
/* Explicit bins by vertex id (LPA_BIN_*); a bin whose kernel cannot hold a
 * vertex's degree is promoted to the next one that can. */
int lpa_schedule_bins(lpa_engine *e, const uint8_t *bins, uint32_t relayout);
/* Vertices per bin kind after scheduling (LPA_NBIN_KINDS entries). */
int lpa_schedule_counts(lpa_engine *e, uint64_t *counts);
/* Per bin kind, 4 words: vertices, arcs, CGR stream bits, arcs read from
 * the small-vertex CSR (4*LPA_NBIN_KINDS entries); the algorithmic bytes of
 * each kernel family follow from these. */
int lpa_schedule_stats(lpa_engine *e, uint64_t *stats);

/* labels: n*K entries, layout labels[v*K + k]; NULL -> L0(v) = v for every k.
 * In semi mode entries != 0xFFFFFFFF are clamped seeds. */
int lpa_labels_init(lpa_engine *e, const uint32_t *labels);
/* Current labels (n*K). */
int lpa_labels_get(lpa_engine *e, uint32_t *labels);

/* One synchronous fused decode+LP step over every bin. */
int lpa_nova_iterate(lpa_engine *e, uint64_t *changed);
/* Decode the compressed part to a CSR on the device (CGLP kernel 1). */
int lpa_cglp_decode(lpa_engine *e);
/* One synchronous step the two-kernel way: decode to CSR, then LP over it. */
int lpa_cglp_iterate(lpa_engine *e, uint64_t *changed);
/* One synchronous step of LP over the uncompressed CSR (the compressed part
 * decoded once, then reused; no per-step decode): the comparator that shows
 * what reading the compressed stream costs.  Same labels as the other two. */
int lpa_csr_iterate(lpa_engine *e, uint64_t *changed);
/* Copy of the device CSR by vertex id: row (n+1), col (m). */
int lpa_cglp_csr_get(lpa_engine *e, uint64_t *row, uint32_t *col, uint64_t *m);

/* ---- learned schedule (no reference counterpart) ----
 * A per-vertex work recorder and an 8 -> 32 -> 32 -> 3 MLP cost model
 * trained on the device.  The model predicts log2(1 + lane-cycles) of a
 * vertex under each kernel family (0: wave-level kernels, 1: workgroup
 * kernel with an LDS table, 2: chunked heavy-vertex kernels); the learned
 * schedule bins each vertex into the family of least predicted cost.  It
 * never changes results and never moves vertices in or out of CMS. */
typedef struct lpa_mlp lpa_mlp;
#define LPA_MLP_NPARAM 1443
#define LPA_NFEAT 8

/* Model inputs by vertex id (n x 8): log2(1 + x) of degree, #intervals,
 * #residuals, encoded bits, #interval segments, #residual segments, mean
 * interval length, and the previous-iteration change flag of label vector 0. */
int lpa_model_features(lpa_engine *e, float *x);
/* Runs one synchronous step with every vertex the schedule allows forced
 * into kernel family `family` (0..2) and the recorder on: cost[v] receives
 * v's share of the device time in picoseconds (each bin's HIP-event time
 * split among its vertices in proportion to the lane-cycles spent on
 * them), measured[v] = 1 when v ran in that family.  Labels and the
 * previous schedule are restored afterwards. */
int lpa_record_family(lpa_engine *e, int family, uint64_t *cost, uint8_t *measured);

int lpa_mlp_create(int device, uint64_t seed, lpa_mlp **out);
int lpa_mlp_destroy(lpa_mlp *m);
int lpa_mlp_get_params(lpa_mlp *m, float *params);
int lpa_mlp_set_params(lpa_mlp *m, const float *params);
/* y (n x 3) = model(x (n x 8)); host arrays. */
int lpa_mlp_forward(lpa_mlp *m, const float *x, uint64_t n, float *y);
/* Mean over samples of the masked squared error and its gradient (1443). */
int lpa_mlp_loss_grad(lpa_mlp *m, const float *x, const float *target, const uint8_t *mask, uint64_t n,
                      float *loss, float *grad);
/* Adam training on the device: epochs over shuffled mini-batches; loss of
 * each epoch (mean over its batches) in loss_per_epoch. */
int lpa_mlp_train(lpa_mlp *m, const float *x, const float *target, const uint8_t *mask, uint64_t n,
                  int epochs, int batch, float lr, uint64_t seed, float *loss_per_epoch);
/* Bins every vertex by the model (argmin of the predicted cost over the
 * families that can hold its degree; CMS vertices stay CMS) and lays the
 * graph out accordingly.  A vertex leaves the family of the default degree
 * schedule only when another is predicted cheaper by more than ~41 %, and
 * the wave family is offered up to degree 512; vertices of the small CSR
 * part keep their degree family. */
int lpa_schedule_learned(lpa_engine *e, lpa_mlp *m);
/* The learned schedule judged by the degree schedule it replaces: the
 * degree layout, the learned layout and the degree layout again each run
 * `iters` fused steps from the current labels (put back afterwards); the
 * learned layout is kept unless the degree layout's median step (over both
 * of its runs) is faster, in which case the degree layout stays (fallback).
 * Reports both medians (ms) and which was kept. */
int lpa_schedule_learned_judged(lpa_engine *e, lpa_mlp *m, int iters, float *learned_ms, float *degree_ms,
                                int *kept_learned);

/* Runs iters steps (method 0 = NOVA fused, 1 = CGLP, 2 = uncompressed CSR) and reports the
 * HIP-event time of each (ms) and the changed counts; captures the step in
 * a HIP graph when use_graph != 0. */
int lpa_time_iterations(lpa_engine *e, int method, int iters, int use_graph,
                        float *ms, uint64_t *changed);
/* Replays steps from HIP graphs captured once per label-buffer parity
 * (on != 0), or launches them eagerly (0, the default).  Refused
 * (LPA_ERR_STATE) while the stream is host-resident. */
int lpa_set_graph_mode(lpa_engine *e, int on);

/* where the CGR stream lives (lpa_stream_residence) */
enum {
    LPA_STREAM_DEVICE = 0,  /* in HBM (the default) */
    LPA_STREAM_HOST = 1     /* in pinned host memory, streamed to the device every step */
};
/* Out-of-core operation for graphs whose CGR stream does not fit in HBM next
 * to the labels.  With LPA_STREAM_HOST the stream (as uploaded and as
 * scheduled) is kept in pinned host memory; every fused or two-kernel step
 * copies the scheduled stream to the device in batches of whole records of
 * at most batch_bytes (0: 1 GiB; a longer record is a batch of its own) into
 * two HBM staging buffers, copying batch i+1 on a second stream while batch
 * i's kernels run.  Offsets, segment metadata, labels, the heavy vertices'
 * decoded lists and the small-vertex CSR stay in HBM; results are those of
 * the HBM-resident stream.  Steps are not captured in graphs meanwhile.  May
 * be called before or after uploading and scheduling: it moves the stream
 * and plans the batches of the current layout (every later layout is
 * planned in turn).  Replaces OutOfCore (include/outofcore.cuh:9-96:
 * compute_batch_boundaries, transfer_batch, transfer_next_batch), which
 * batches the neighbour lists of an uncompressed CSR the same way. */
int lpa_stream_residence(lpa_engine *e, int where, uint64_t batch_bytes);
/* Batches of the current layout and the bytes of both staging buffers
 * (0 and 0 while the stream is in HBM). */
int lpa_stream_batches(lpa_engine *e, uint64_t *nbatches, uint64_t *stage_bytes);
/* Times `iters` fused steps for each candidate schedule (lane, wave and
 * quarter-wave thresholds, then the bin order and XCD tile mapping for the
 * best thresholds), keeps the fastest (median), and restores the labels
 * after every candidate, so tuning never changes results. */
int lpa_autotune(lpa_engine *e, int iters, lpa_sched_params *best, float *best_ms);
/* Average time (ms) per step of each kernel family in the last eager timed
 * run (HIP events on the engine stream). */
int lpa_last_kernel_times(lpa_engine *e, float *ms_per_kind /* LPA_NBIN_KINDS */);
/* Measurement helper, no counterpart in the reference: copies `bytes`
 * (a multiple of 16) between two fresh device buffers `reps` times with a
 * plain 16-byte-per-lane copy kernel (k_calibration_copy), so hardware
 * counters of the step kernels can be scaled by what they read for a copy
 * of known size; *ms = average kernel time. */
int lpa_calibration_copy(lpa_engine *e, uint64_t bytes, int reps, float *ms);
/* Measurement helper, no counterpart in the reference: random row gathers,
 * the access pattern of a step's label reads.  `gathers` reads of rows of
 * row_words 32-bit words (1, 2, 4, 8 or 16) at uniformly random rows of a
 * fresh device buffer of `bytes` (k_calibration_gather, every read
 * independent of the others), `reps` times; *ms = average kernel time, so
 * gathers / ms is the rate the memory system sustains for this pattern at
 * this footprint. */
int lpa_calibration_gather(lpa_engine *e, uint64_t bytes, uint32_t row_words, uint64_t gathers, int reps,
                           float *ms);

#ifdef __cplusplus
}
#endif
#endif
