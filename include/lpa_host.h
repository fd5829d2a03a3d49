/*
 * lpa_host.h -- C ABI of the host side of the MI355X label-propagation engine
 * (liblpa_host.so, plain C++17 + OpenMP): graph generators, CSR building,
 * BFS relabelling, the CGR encoder and the reference's on-disk file set.
 *
 * Reference interfaces these replace:
 *   lpah_cgr_encode      CgrCompressor::compress + write_bit_array
 *                        (src/compressor.cpp:70-100, 178-290; include/compressor.h:23-56)
 *   lpah_files_write     CgrCompressor::write_cgr + File::write_binary
 *                        (src/compressor.cpp:157-176; src/file.cpp:76-95)
 *   lpah_files_load      File::load_graph / load_compress_graph
 *                        (src/file.cpp:24-74, 197-262; include/file.h:28-76)
 *   lpah_csr_from_edges  File::read_from_edge_list / CgrCompressor::load_graph
 *                        (src/file.cpp:115-195; src/compressor.cpp:22-68)
 *   generators, BFS      no reference counterpart (synthetic inputs)
 *
 * Conventions: every function returns 0 (LPA_OK) or a negative status; the
 * message is available from lpah_last_error() (thread-local).  Objects are
 * opaque handles owned by the library until the matching *_free call.
 * Vertex ids are uint32 (< 2^30), arc offsets and bit offsets uint64.
 */
#ifndef LPA_HOST_H
#define LPA_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lpah_csr lpah_csr;   /* CSR graph: row[n+1] (u64), col[m] (u32), rows sorted, no duplicates */
typedef struct lpah_cgr lpah_cgr;   /* CGR file set in memory */

const char *lpah_last_error(void);

/* ---- CSR ---- */
/* Edge list -> CSR.  undirected != 0 adds the reverse arc of every edge
 * (the reference's one_way flag, file.cpp:136-140).  Self loops and
 * duplicate arcs are removed (the reference breaks on them: a duplicate
 * gives a gap of -1, compressor.cpp:459). */
int lpah_csr_from_edges(uint64_t n, const uint32_t *src, const uint32_t *dst, uint64_t m,
                        int undirected, lpah_csr **out);
/* Wrap caller arrays (copied). */
int lpah_csr_from_arrays(uint64_t n, const uint64_t *row, const uint32_t *col, lpah_csr **out);
/* Read a text edge list "u v" per line (file.cpp:115-195). */
int lpah_csr_read_edge_list(const char *path, int undirected, lpah_csr **out);
uint64_t lpah_csr_n(const lpah_csr *g);
uint64_t lpah_csr_m(const lpah_csr *g);
const uint64_t *lpah_csr_row(const lpah_csr *g);
const uint32_t *lpah_csr_col(const lpah_csr *g);
uint32_t lpah_csr_max_degree(const lpah_csr *g);
void lpah_csr_free(lpah_csr *g);

/* Relabel vertices in BFS order from `root` (neighbours visited in
 * ascending id order; unreached vertices appended in ascending id order),
 * then re-sort every row.  perm_out (n entries, may be NULL) gets new_id[old]. */
int lpah_csr_bfs_relabel(const lpah_csr *g, uint32_t root, uint32_t *perm_out, lpah_csr **out);

/* ---- deterministic synthetic graphs (counter-based splitmix64 streams) ---- */
/* side x side 4-neighbour lattice, id = r*side + c; each lattice edge dropped
 * with probability p_drop; floor(shortcut_frac * side^2) distinct random
 * shortcut edges added; undirected. */
int lpah_gen_lattice(uint32_t side, double p_drop, double shortcut_frac, uint64_t seed,
                     lpah_csr **out);
/* R-MAT: edge_factor * 2^scale draws with quadrant probabilities a,b,c
 * (d = 1-a-b-c); symmetrised, self loops and duplicates removed; BFS
 * relabelled from vertex 0 when bfs != 0. */
int lpah_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                  uint64_t seed, int bfs, lpah_csr **out);
/* Directed copying model: n vertices; out-degree from a power law with the
 * given exponent, capped, x_min chosen so the expected mean is mean_deg;
 * each out-arc copies (with probability copy_frac) a random target of a
 * prototype drawn uniformly within +-window ids, else a uniform target. */
int lpah_gen_copying(uint64_t n, double mean_deg, double exponent, uint32_t cap,
                     double copy_frac, uint32_t window, uint64_t seed, lpah_csr **out);

/* ---- CGR ---- */
/* Encode with the reference's parameters (zeta 3, min interval 4, 256-bit
 * segments): vertices with deg >= large_thresh first (by id), then
 * small_thresh <= deg < large_thresh (by id), then deg < small_thresh go
 * to the small CSR (by id). */
int lpah_cgr_encode(const lpah_csr *g, uint32_t small_thresh, uint32_t large_thresh,
                    lpah_cgr **out);
uint64_t lpah_cgr_n(const lpah_cgr *c);          /* all vertices */
uint64_t lpah_cgr_nc(const lpah_cgr *c);         /* compressed vertices */
uint64_t lpah_cgr_nbits(const lpah_cgr *c);
/* big-endian 32-bit words as File::graph holds them (file.cpp:222-238),
 * followed by 4 zero padding words; lpah_cgr_nwords counts the data words */
const uint32_t *lpah_cgr_words(const lpah_cgr *c);
uint64_t lpah_cgr_nwords(const lpah_cgr *c);
const uint64_t *lpah_cgr_offsets(const lpah_cgr *c);   /* n_c + 1 START bit offsets */
const uint32_t *lpah_cgr_vertices(const lpah_cgr *c);  /* n */
const uint64_t *lpah_cgr_small_offsets(const lpah_cgr *c); /* n - n_c + 1 */
const uint32_t *lpah_cgr_small_nbrs(const lpah_cgr *c);
uint64_t lpah_cgr_small_m(const lpah_cgr *c);
void lpah_cgr_free(lpah_cgr *c);

/* Write <path>.graph/.offset/.vertices.bin/[.graph.small.bin/.offset.small.bin]
 * byte-identical to the reference writer, plus <path>.row/.clm (uint32 CSR,
 * skipped when they do not fit 32 bits) when csr != NULL, plus the sidecar
 * <path>.offset64.bin (uint64 start offsets) that large graphs need. */
int lpah_files_write(const lpah_cgr *c, const lpah_csr *csr, const char *path);
/* Load a file set written by the reference or by lpah_files_write.
 * small_thresh > 0 loads the small CSR (the reference's has_small). */
int lpah_files_load(const char *path, int small_thresh, lpah_cgr **out);

/* Number of OpenMP threads used. */
int lpah_threads(void);

#ifdef __cplusplus
}
#endif
#endif
