/*
 * lpa_oracle.h -- CPU oracle for the CGR label-propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it.  The product (lpa_amd/) never links, loads or calls it.
 *
 * It restates, in plain C:
 *   - the CGR bitstream written by the reference compressor
 *       (src/compressor.cpp:122-155 codes, :306-515 vertex records,
 *        :178-290 file layout; include/compressor.h:66-99 code sizes),
 *   - the decoder the reference runs on the device
 *       (include/graph_utils.cuh:19-183; segment stride 256 and where the
 *        residual block starts: include/decoder.cuh:139-163),
 *   - synchronous label propagation with a smallest-label tie-break
 *     ("LP-SC"): exact mode, count-min-sketch mode, K batched label
 *     vectors and semi-supervised clamping.  The reference has no CPU LP;
 *     the semantics are the task's (the reference kernels are asynchronous
 *     and prefer larger labels, baselines/nova/kernel.cuh:117-122).
 *
 * Pinning (tests/test_oracle_*.py): golden file sets written by the
 * reference compressor itself (built from its sources under oracle/_ref),
 * the hand-derived record of vertex 5 with N = {1,6,7,8,9,12}, closed forms
 * of the gamma/zeta code lengths, brute-force mode counting, and closed-form
 * LP trajectories on rings, stars, complete graphs and paths.
 *
 * Every function returns >= 0 on success and a negative code on malformed
 * input (out-of-range bit reads, neighbour ids >= n, buffer too small).
 */
#ifndef LPA_ORACLE_H
#define LPA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_UNLABELLED 0xFFFFFFFFu

/* Elias-gamma length of x >= 0: 2*floor(log2(x+1)) + 1
 * (compressor.h:71-79, compressor.cpp:122-129). */
uint32_t orc_gamma_len(uint64_t x);
/* zeta_3 length of x >= 0: (floor(floor(log2(x+1))/3) + 1) * 4
 * (compressor.h:81-90, compressor.cpp:131-146). */
uint32_t orc_zeta_len(uint64_t x);

/* Bit length of the record of every vertex of a CSR graph whose rows are
 * sorted ascending and free of duplicates (compressor.cpp:306-515 with the
 * parameters File uses: zeta 3, min interval 4, 256-bit segments,
 * file.cpp:84).  rec_bits has n entries. */
int orc_record_bits(uint64_t n, const uint64_t *row, const uint32_t *col,
                    uint64_t *rec_bits);

/* Cost-model features, 8 floats per vertex (no reference counterpart):
 * degree, #intervals, #residuals, record bits, #interval segments,
 * #residual segments, mean interval length, log2(1 + degree); vertices of
 * degree < small_thresh (small CSR) get {d, 0, d, 32 d, 0, 0, 0, log2(1+d)}. */
int orc_record_features(uint64_t n, const uint64_t *row, const uint32_t *col,
                        uint32_t small_thresh, float *feat);

/* Write the reference file set in memory (compressor.cpp:178-290):
 *   stream      MSB-first bytes, ceil(nbits/8) of them (caller zeroes it),
 *   end_offs    n_c cumulative END bit offsets (the .offset lines after the count),
 *   vertices    n ids: deg>=large by id, then small<=deg<large by id, then deg<small by id,
 *   small_offs  n_s+1 CSR offsets of the deg<small vertices (.offset.small.bin),
 *   small_nbrs  their neighbours (.graph.small.bin).
 * out3 receives {n_c, nbits, m_s}. */
int orc_write_cgr(uint64_t n, const uint64_t *row, const uint32_t *col,
                  uint32_t small_thresh, uint32_t large_thresh,
                  uint8_t *stream, uint64_t stream_cap,
                  uint64_t *end_offs, uint32_t *vertices,
                  uint32_t *small_offs, uint32_t *small_nbrs,
                  uint64_t *out3);

/* Decode the record of vertex v starting at bit pos of an MSB-first byte
 * stream of nbits bits (graph_utils.cuh:19-183).  Writes up to cap
 * neighbours to out in encoded order; returns the degree, or < 0 if the
 * record reads past nbits.  *end_pos (may be NULL) receives the bit just
 * after the record. */
int64_t orc_decode_record(const uint8_t *stream, uint64_t nbits, uint64_t pos,
                          uint32_t v, uint32_t *out, uint64_t cap,
                          uint64_t *end_pos);

/* Multiply-shift row hash of the count-min sketch: a_r, b_r drawn from
 * splitmix64 seeded with seed (a_r forced odd); h_r(x) = (a_r*x + b_r) >> (64-log2 w).
 * Writes 2*depth words {a_0,b_0,a_1,b_1,...}. */
void orc_cms_params(uint64_t seed, uint32_t depth, uint64_t *ab);

/* LP configuration shared by the LP entry points. */
typedef struct {
    uint32_t K;           /* label vectors, layout labels[v*K + k] */
    int32_t semi;         /* 1: UNLABELLED neighbours ignored, seeds clamped */
    uint32_t cms_thresh;  /* 0: exact mode everywhere; else CMS when deg > cms_thresh */
    uint32_t cms_depth;
    uint32_t cms_width;   /* power of two */
    uint64_t cms_seed;
    int32_t nthreads;     /* OpenMP threads, <= 0: library default */
} orc_lp_cfg;

/* One synchronous step over a CSR graph: out[v*K+k] = f(labels of N(v)).
 * seeds (n*K bytes, may be NULL) marks clamped entries in semi mode.
 * Returns the number of (v,k) entries that changed. */
int64_t orc_step_csr(uint64_t n, const uint64_t *row, const uint32_t *col,
                     const orc_lp_cfg *cfg, const uint8_t *seeds,
                     const uint32_t *in, uint32_t *out);

/* One synchronous step that decodes the CGR file set every time, as the
 * hot path does: positions 0..n_c-1 are CGR records starting at offs[i]
 * (n_c+1 START offsets, offs[0]=0), positions n_c..n-1 come from the small
 * CSR.  vertices[i] names the vertex at position i.  Returns the change
 * count, or < 0 on a malformed record. */
int64_t orc_step_cgr(uint64_t n, const uint8_t *stream, uint64_t nbits,
                     uint64_t n_c, const uint64_t *offs, const uint32_t *vertices,
                     const uint64_t *small_offs, const uint32_t *small_nbrs,
                     const orc_lp_cfg *cfg, const uint8_t *seeds,
                     const uint32_t *in, uint32_t *out);

/* Same as orc_step_cgr restricted to the positions listed in pos[0..np);
 * entries of out outside those vertices are left untouched.  Used for the
 * one-step check of large runs. */
int64_t orc_step_cgr_subset(uint64_t n, const uint8_t *stream, uint64_t nbits,
                            uint64_t n_c, const uint64_t *offs, const uint32_t *vertices,
                            const uint64_t *small_offs, const uint32_t *small_nbrs,
                            const orc_lp_cfg *cfg, const uint8_t *seeds,
                            const uint64_t *pos, uint64_t np,
                            const uint32_t *in, uint32_t *out);

/* Run LP from labels (in/out, n*K).  Iterates until nothing changes (when
 * until_converged) or max_iters steps.  changed[it] receives each step's
 * change count; returns the number of steps run, < 0 on error.
 * Uses CGR when stream != NULL, else the CSR (row, col). */
int64_t orc_lp_run(uint64_t n, const uint64_t *row, const uint32_t *col,
                   const uint8_t *stream, uint64_t nbits, uint64_t n_c,
                   const uint64_t *offs, const uint32_t *vertices,
                   const uint64_t *small_offs, const uint32_t *small_nbrs,
                   const orc_lp_cfg *cfg, uint32_t *labels,
                   int32_t max_iters, int32_t until_converged, uint64_t *changed);

/* Number of OpenMP threads the library will use by default. */
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
