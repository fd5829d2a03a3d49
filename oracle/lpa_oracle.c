/*
 * lpa_oracle.c -- CPU oracle (test infrastructure only; see lpa_oracle.h).
 *
 * Written for obviousness, not speed: the encoder keeps one byte per bit,
 * the decoder reads codes through a bounds-checked 64-bit window, and the
 * mode is found by sorting each neighbourhood's labels (radix sort).
 */
#include "lpa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SEG_BITS 256u   /* 8*32 bits per segment (file.cpp:84, graph_utils.cuh:17) */
#define MIN_ITV 4u      /* minimum interval length (file.cpp:84, graph_utils.cuh:172) */

/* ---------------------------------------------------------------- codes */

static int floor_log2(uint64_t y) { /* y > 0 */
    int r = 0;
    while (y > 1) { y >>= 1; r++; }
    return r;
}

/* compressor.h:71-79 */
uint32_t orc_gamma_len(uint64_t x) { return 2u * (uint32_t)floor_log2(x + 1) + 1u; }

/* compressor.h:81-90, zeta_k = 3 */
uint32_t orc_zeta_len(uint64_t x) { return ((uint32_t)floor_log2(x + 1) / 3u + 1u) * 4u; }

/* int_2_nat, compressor.h:66-69 */
static uint64_t to_nat(int64_t d) { return d >= 0 ? (uint64_t)d << 1 : (uint64_t)(-2 * d - 1); }

/* --------------------------------------------------------- bit vectors */

typedef struct { uint8_t *b; uint64_t len, cap; } bitv;

static int bv_reserve(bitv *v, uint64_t need) {
    if (need <= v->cap) return 0;
    uint64_t c = v->cap ? v->cap : 64;
    while (c < need) c *= 2;
    uint8_t *p = (uint8_t *)realloc(v->b, c);
    if (!p) return -1;
    v->b = p; v->cap = c;
    return 0;
}
/* encode(): x in len bits, MSB first (compressor.cpp:148-155) */
static int bv_put(bitv *v, uint64_t x, int len) {
    if (bv_reserve(v, v->len + (uint64_t)len)) return -1;
    for (int i = len - 1; i >= 0; i--) v->b[v->len++] = (uint8_t)((x >> i) & 1u);
    return 0;
}
static int bv_cat(bitv *v, const bitv *w) {
    if (bv_reserve(v, v->len + w->len)) return -1;
    if (w->len) memcpy(v->b + v->len, w->b, w->len);
    v->len += w->len;
    return 0;
}
/* encode_gamma, compressor.cpp:122-129 */
static int bv_gamma(bitv *v, uint64_t x) {
    x++;
    int L = floor_log2(x);
    if (bv_put(v, 1, L + 1)) return -1;
    return bv_put(v, x, L);
}
/* encode_zeta, compressor.cpp:131-146 (k = 3) */
static int bv_zeta(bitv *v, uint64_t x) {
    x++;
    int h = floor_log2(x) / 3;
    if (bv_put(v, 1, h + 1)) return -1;
    return bv_put(v, x, (h + 1) * 3);
}

/* ------------------------------------------------------------- encoder */

typedef struct { uint64_t cnt; bitv bits; } seg_t;
typedef struct { seg_t *s; uint64_t n, cap; } segvec;

static int segs_push(segvec *sv, uint64_t cnt, bitv *bits) {
    if (sv->n == sv->cap) {
        uint64_t c = sv->cap ? 2 * sv->cap : 8;
        seg_t *p = (seg_t *)realloc(sv->s, c * sizeof(seg_t));
        if (!p) return -1;
        sv->s = p; sv->cap = c;
    }
    sv->s[sv->n].cnt = cnt;
    sv->s[sv->n].bits = *bits;   /* takes ownership */
    sv->n++;
    bits->b = NULL; bits->len = bits->cap = 0;
    return 0;
}
static void segs_free(segvec *sv) {
    for (uint64_t i = 0; i < sv->n; i++) free(sv->s[i].bits.b);
    free(sv->s);
}

/* append_segment, compressor.cpp:501-515 */
static int put_segment(bitv *rec, uint64_t cnt, const bitv *body, uint64_t align) {
    uint64_t start = rec->len;
    if (bv_gamma(rec, cnt) || bv_cat(rec, body)) return -1;
    if (align && rec->len - start > align) return -2;
    while (rec->len - start < align)
        if (bv_put(rec, 0, 1)) return -1;
    return 0;
}

/* encode_node (compressor.cpp:306-328) with intervalize (:330-362),
 * encode_intervals (:364-439) and encode_residuals (:441-499). */
/* what encode_record saw: #intervals, #residuals, #interval segments,
 * #residual segments, sum of interval lengths (the cost-model features) */
typedef struct { uint64_t ni, nr, si, sr, itv_sum; } rec_stats;

static int encode_record(uint64_t v, const uint32_t *adj, uint64_t d, bitv *rec, rec_stats *st) {
    int rc = -1;
    uint64_t ni = 0, nr = 0;
    uint64_t *il = (uint64_t *)malloc((d / MIN_ITV + 1) * sizeof(uint64_t));
    uint64_t *ilen = (uint64_t *)malloc((d / MIN_ITV + 1) * sizeof(uint64_t));
    uint64_t *res = (uint64_t *)malloc((d + 1) * sizeof(uint64_t));
    segvec segs = {0};
    bitv cur = {0};
    if (!il || !ilen || !res) goto out;

    /* intervalize: maximal +1 runs; len >= 4 -> interval, else residuals */
    for (uint64_t a = 0; a < d;) {
        uint64_t b = a + 1;
        while (b < d && (uint64_t)adj[b - 1] + 1 == adj[b]) b++;
        if (b - a >= MIN_ITV) { il[ni] = adj[a]; ilen[ni] = b - a; ni++; }
        else for (uint64_t t = a; t < b; t++) res[nr++] = adj[t];
        a = b;
    }

    /* interval block */
    uint64_t cnt = 0;
    for (uint64_t i = 0; i < ni; i++) {
        uint64_t code_left = cnt == 0 ? to_nat((int64_t)il[i] - (int64_t)v)
                                      : il[i] - il[i - 1] - ilen[i - 1] - 1;
        uint64_t code_len = ilen[i] - MIN_ITV;
        if (orc_gamma_len(cnt + 1) + cur.len + orc_gamma_len(code_left) + orc_gamma_len(code_len) > SEG_BITS) {
            if (segs_push(&segs, cnt, &cur)) goto out;
            cnt = 0;
            code_left = to_nat((int64_t)il[i] - (int64_t)v);
        }
        cnt++;
        if (bv_gamma(&cur, code_left) || bv_gamma(&cur, code_len)) goto out;
    }
    if (segs.n == 0) {
        if (segs_push(&segs, cnt, &cur)) goto out;
    } else {
        /* the trailing partial segment is merged into the last full one,
         * continuing its gap coding (compressor.cpp:412-425) */
        seg_t *last = &segs.s[segs.n - 1];
        last->cnt += cnt;
        for (uint64_t i = ni - cnt; i < ni; i++)
            if (bv_gamma(&last->bits, il[i] - il[i - 1] - ilen[i - 1] - 1) ||
                bv_gamma(&last->bits, ilen[i] - MIN_ITV)) goto out;
        free(cur.b); cur.b = NULL; cur.len = cur.cap = 0;
    }
    if (st) {
        st->ni = ni;
        st->nr = nr;
        st->si = segs.n;
        st->itv_sum = 0;
        for (uint64_t i = 0; i < ni; i++) st->itv_sum += ilen[i];
    }
    if (bv_gamma(rec, segs.n - 1)) goto out;
    for (uint64_t i = 0; i < segs.n; i++)
        if (put_segment(rec, segs.s[i].cnt, &segs.s[i].bits, i + 1 == segs.n ? 0 : SEG_BITS)) goto out;
    segs_free(&segs);
    memset(&segs, 0, sizeof segs);

    /* residual block */
    cnt = 0;
    for (uint64_t i = 0; i < nr; i++) {
        uint64_t code = cnt == 0 ? to_nat((int64_t)res[i] - (int64_t)v) : res[i] - res[i - 1] - 1;
        if (orc_gamma_len(cnt + 1) + cur.len + orc_zeta_len(code) > SEG_BITS) {
            if (segs_push(&segs, cnt, &cur)) goto out;
            cnt = 0;
            code = to_nat((int64_t)res[i] - (int64_t)v);
        }
        cnt++;
        if (bv_zeta(&cur, code)) goto out;
    }
    if (segs.n == 0) {
        if (segs_push(&segs, cnt, &cur)) goto out;
    } else {
        seg_t *last = &segs.s[segs.n - 1];
        last->cnt += cnt;
        for (uint64_t i = nr - cnt; i < nr; i++)
            if (bv_zeta(&last->bits, res[i] - res[i - 1] - 1)) goto out;
    }
    if (st) st->sr = segs.n;
    if (bv_gamma(rec, segs.n - 1)) goto out;
    for (uint64_t i = 0; i < segs.n; i++)
        if (put_segment(rec, segs.s[i].cnt, &segs.s[i].bits, i + 1 == segs.n ? 0 : SEG_BITS)) goto out;
    rc = 0;
out:
    segs_free(&segs);
    free(cur.b);
    free(il); free(ilen); free(res);
    return rc;
}

static int check_rows(uint64_t n, const uint64_t *row, const uint32_t *col) {
    for (uint64_t v = 0; v < n; v++) {
        if (row[v + 1] < row[v]) return -3;
        for (uint64_t e = row[v]; e < row[v + 1]; e++) {
            if (col[e] >= n) return -4;
            if (e > row[v] && col[e] <= col[e - 1]) return -5;   /* unsorted or duplicate */
        }
    }
    return 0;
}

int orc_record_bits(uint64_t n, const uint64_t *row, const uint32_t *col, uint64_t *rec_bits) {
    int rc = check_rows(n, row, col);
    if (rc) return rc;
    bitv rec = {0};
    for (uint64_t v = 0; v < n; v++) {
        rec.len = 0;
        if (encode_record(v, col + row[v], row[v + 1] - row[v], &rec, NULL)) { free(rec.b); return -1; }
        rec_bits[v] = rec.len;
    }
    free(rec.b);
    return 0;
}

/* Cost-model features of every vertex (8 floats each), as the engine
 * reports them for its compressed vertices: degree, #intervals, #residuals,
 * record bits, #interval segments, #residual segments, mean interval length
 * (0 without intervals), log2(1 + degree).  A vertex of degree < small_thresh
 * is read from the small CSR: {d, 0, d, 32 d, 0, 0, 0, log2(1 + d)}.  The
 * counts are those of the record encode_record writes; no reference
 * counterpart (the reference has no cost model). */
int orc_record_features(uint64_t n, const uint64_t *row, const uint32_t *col, uint32_t small_thresh,
                        float *feat) {
    int rc = check_rows(n, row, col);
    if (rc) return rc;
    bitv rec = {0};
    for (uint64_t v = 0; v < n; v++) {
        const uint64_t d = row[v + 1] - row[v];
        float *f = feat + v * 8;
        if (d < small_thresh) {
            f[0] = (float)d; f[1] = 0.0f; f[2] = (float)d; f[3] = 32.0f * (float)d;
            f[4] = f[5] = f[6] = 0.0f;
            f[7] = log2f(1.0f + (float)d);
            continue;
        }
        rec_stats st = {0};
        rec.len = 0;
        if (encode_record(v, col + row[v], d, &rec, &st)) { free(rec.b); return -1; }
        f[0] = (float)d;
        f[1] = (float)st.ni;
        f[2] = (float)st.nr;
        f[3] = (float)rec.len;
        f[4] = (float)st.si;
        f[5] = (float)st.sr;
        f[6] = st.ni ? (float)st.itv_sum / (float)st.ni : 0.0f;
        f[7] = log2f(1.0f + (float)d);
    }
    free(rec.b);
    return 0;
}

/* write_bit_array, compressor.cpp:178-290 */
int orc_write_cgr(uint64_t n, const uint64_t *row, const uint32_t *col,
                  uint32_t small_thresh, uint32_t large_thresh,
                  uint8_t *stream, uint64_t stream_cap,
                  uint64_t *end_offs, uint32_t *vertices,
                  uint32_t *small_offs, uint32_t *small_nbrs,
                  uint64_t *out3) {
    int rc = check_rows(n, row, col);
    if (rc) return rc;
    bitv rec = {0};
    uint64_t nbits = 0, nc = 0, np = 0;
    for (int pass = 0; pass < 2; pass++) {
        for (uint64_t v = 0; v < n; v++) {
            uint64_t d = row[v + 1] - row[v];
            int take = pass == 0 ? d >= large_thresh : (d >= small_thresh && d < large_thresh);
            if (!take) continue;
            rec.len = 0;
            if (encode_record(v, col + row[v], d, &rec, NULL)) { free(rec.b); return -1; }
            for (uint64_t i = 0; i < rec.len; i++, nbits++) {
                if ((nbits >> 3) >= stream_cap) { free(rec.b); return -6; }
                if (rec.b[i]) stream[nbits >> 3] |= (uint8_t)(0x80u >> (nbits & 7));
            }
            end_offs[nc++] = nbits;
            vertices[np++] = (uint32_t)v;
        }
    }
    free(rec.b);
    uint64_t ms = 0, ns = 0;
    small_offs[0] = 0;
    for (uint64_t v = 0; v < n; v++) {
        uint64_t d = row[v + 1] - row[v];
        if (d >= small_thresh) continue;
        for (uint64_t e = row[v]; e < row[v + 1]; e++) small_nbrs[ms++] = col[e];
        vertices[np++] = (uint32_t)v;
        small_offs[++ns] = (uint32_t)ms;
    }
    out3[0] = nc; out3[1] = nbits; out3[2] = ms;
    return 0;
}

/* ------------------------------------------------------------- decoder */

typedef struct { const uint8_t *s; uint64_t nbits; uint64_t pos; int bad; } rdr;

static int rd_bit(rdr *r) {
    if (r->pos >= r->nbits) { r->bad = 1; return 0; }
    int b = (r->s[r->pos >> 3] >> (7 - (r->pos & 7))) & 1;
    r->pos++;
    return b;
}
static uint64_t rd_bits(rdr *r, int len) {
    uint64_t x = 0;
    for (int i = 0; i < len; i++) x = (x << 1) | (uint64_t)rd_bit(r);
    return x;
}
/* unary prefix: number of 0 bits before the next 1; the 1 is not consumed */
static int rd_zeros(rdr *r) {
    int z = 0;
    for (;;) {
        if (r->pos >= r->nbits || z > 62) { r->bad = 1; return 0; }
        if ((r->s[r->pos >> 3] >> (7 - (r->pos & 7))) & 1) return z;
        r->pos++; z++;
    }
}
/* decode_gamma, graph_utils.cuh:48-67 */
static uint64_t rd_gamma(rdr *r) {
    int L = rd_zeros(r);
    return rd_bits(r, L + 1) - 1;
}
/* decode_zeta, graph_utils.cuh:70-77 */
static uint64_t rd_zeta(rdr *r) {
    int h = rd_zeros(r);
    r->pos++;                               /* the terminating 1 */
    return rd_bits(r, (h + 1) * 3) - 1;
}
/* decode_first_num, graph_utils.cuh:34-37 */
static int64_t from_nat(uint32_t v, uint64_t x) {
    return (x & 1) ? (int64_t)v - (int64_t)(x >> 1) - 1 : (int64_t)v + (int64_t)(x >> 1);
}

int64_t orc_decode_record(const uint8_t *stream, uint64_t nbits, uint64_t pos,
                          uint32_t v, uint32_t *out, uint64_t cap, uint64_t *end_pos) {
    rdr r = {stream, nbits, pos, 0};
    uint64_t d = 0;
    /* interval block: segment count, then segments at base + 256*j
     * (graph_utils.cuh:88-97, 139-183; decoder.cuh:139-163) */
    uint64_t nseg = rd_gamma(&r) + 1;
    if (r.bad || nseg > nbits) return -1;
    uint64_t base = r.pos, block_end = r.pos;
    for (uint64_t j = 0; j < nseg; j++) {
        r.pos = base + SEG_BITS * j;
        uint64_t c = rd_gamma(&r);
        int64_t left = 0, stop = 0;
        for (uint64_t t = 0; t < c && !r.bad; t++) {
            if (t == 0) left = from_nat(v, rd_gamma(&r));
            else left = stop + (int64_t)rd_gamma(&r) + 1;
            int64_t len = (int64_t)rd_gamma(&r) + MIN_ITV;
            if (left < 0 || left + len > (int64_t)0x100000000LL) return -2;
            for (int64_t x = left; x < left + len; x++, d++)
                if (d < cap) out[d] = (uint32_t)x;
            stop = left + len;
        }
        if (r.bad) return -1;
        block_end = r.pos;
    }
    /* residual block starts where the last interval segment ends */
    r.pos = block_end;
    nseg = rd_gamma(&r) + 1;
    if (r.bad || nseg > nbits) return -1;
    base = r.pos;
    for (uint64_t j = 0; j < nseg; j++) {
        r.pos = base + SEG_BITS * j;
        uint64_t c = rd_gamma(&r);
        int64_t cur = 0;
        for (uint64_t t = 0; t < c && !r.bad; t++) {
            if (t == 0) cur = from_nat(v, rd_zeta(&r));
            else cur = cur + (int64_t)rd_zeta(&r) + 1;
            if (cur < 0 || cur >= (int64_t)0x100000000LL) return -2;
            if (d < cap) out[d] = (uint32_t)cur;
            d++;
        }
        if (r.bad) return -1;
        block_end = r.pos;
    }
    if (end_pos) *end_pos = block_end;
    return (int64_t)d;
}

/* ------------------------------------------------------------------ LP */

static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_cms_params(uint64_t seed, uint32_t depth, uint64_t *ab) {
    uint64_t s = seed;
    for (uint32_t r = 0; r < depth; r++) {
        ab[2 * r] = splitmix64(&s) | 1ull;
        ab[2 * r + 1] = splitmix64(&s);
    }
}

/* column of label x in sketch row r: top lw bits of a_r*x + b_r (mod 2^64) */
static uint64_t cms_col(const uint64_t *ab, uint32_t r, uint32_t lw, uint32_t x) {
    uint64_t h = ab[2 * r] * (uint64_t)x + ab[2 * r + 1];
    return lw ? h >> (64 - lw) : 0;
}

/* Ascending sort of n labels: insertion sort for short lists, else an LSD
 * radix sort over bytes (tmp holds n entries); byte passes whose digit is
 * the same in every label are skipped. */
static void sort_u32(uint32_t *a, uint64_t n, uint32_t *tmp) {
    if (n <= 48) {
        for (uint64_t i = 1; i < n; i++) {
            uint32_t x = a[i];
            uint64_t j = i;
            while (j > 0 && a[j - 1] > x) { a[j] = a[j - 1]; j--; }
            a[j] = x;
        }
        return;
    }
    uint32_t or_all = 0, and_all = 0xFFFFFFFFu;
    for (uint64_t i = 0; i < n; i++) { or_all |= a[i]; and_all &= a[i]; }
    uint32_t *src = a, *dst = tmp;
    for (int sh = 0; sh < 32; sh += 8) {
        if ((((or_all ^ and_all) >> sh) & 0xFFu) == 0) continue;   /* constant digit */
        uint64_t cnt[257] = {0};
        for (uint64_t i = 0; i < n; i++) cnt[((src[i] >> sh) & 0xFFu) + 1]++;
        for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFFu]++] = src[i];
        uint32_t *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(uint32_t));
}

typedef struct {
    uint32_t *nbr; uint64_t nbr_cap;
    uint32_t *lab; uint64_t lab_cap;   /* K x d gathered labels, vector-major */
    uint32_t *tmp; uint64_t tmp_cap;   /* radix-sort buffer */
    uint32_t *sketch;
} scratch_t;

static int grow(uint32_t **p, uint64_t *cap, uint64_t need) {
    if (need <= *cap) return 0;
    uint64_t c = *cap ? *cap : 256;
    while (c < need) c *= 2;
    uint32_t *q = (uint32_t *)realloc(*p, c * sizeof(uint32_t));
    if (!q) return -1;
    *p = q; *cap = c;
    return 0;
}

/* f over the labels lab[0..d) of one vertex's neighbours for one vector
 * (LP-SC): most frequent label, ties to the smallest; unlabelled entries
 * ignored; no labelled neighbour -> keep; CMS estimate when d > cms_thresh.
 * lab is reordered. */
static uint32_t mode_of(const orc_lp_cfg *c, const uint64_t *ab, scratch_t *s, uint32_t *lab, uint64_t d,
                        uint32_t old) {
    uint64_t m = 0;
    for (uint64_t j = 0; j < d; j++)
        if (lab[j] != ORC_UNLABELLED) lab[m++] = lab[j];
    if (m == 0) return old;
    if (c->cms_thresh && d > c->cms_thresh) {
        uint32_t w = c->cms_width, lw = 0;
        while ((1u << lw) < w) lw++;
        memset(s->sketch, 0, (size_t)c->cms_depth * w * sizeof(uint32_t));
        for (uint64_t j = 0; j < m; j++)
            for (uint32_t r = 0; r < c->cms_depth; r++)
                s->sketch[(uint64_t)r * w + cms_col(ab, r, lw, lab[j])]++;
        uint32_t best = 0, best_est = 0;
        for (uint64_t j = 0; j < m; j++) {
            uint32_t est = 0xFFFFFFFFu;
            for (uint32_t r = 0; r < c->cms_depth; r++) {
                uint32_t e = s->sketch[(uint64_t)r * w + cms_col(ab, r, lw, lab[j])];
                if (e < est) est = e;
            }
            if (est > best_est || (est == best_est && lab[j] < best)) { best_est = est; best = lab[j]; }
        }
        return best;
    }
    sort_u32(lab, m, s->tmp);
    uint32_t best = lab[0];
    uint64_t best_run = 0, run = 0;
    for (uint64_t j = 0; j < m; j++) {
        run = (j > 0 && lab[j] == lab[j - 1]) ? run + 1 : 1;
        if (run > best_run) { best_run = run; best = lab[j]; }   /* strict: first (smallest) wins */
    }
    return best;
}

typedef struct {
    uint64_t n;
    const uint64_t *row; const uint32_t *col;               /* CSR source */
    const uint8_t *stream; uint64_t nbits; uint64_t n_c;    /* CGR source */
    const uint64_t *offs; const uint32_t *vertices;
    const uint64_t *soffs; const uint32_t *snbrs;
} gview;

/* neighbours of the vertex at position i (CGR) or of vertex i (CSR) */
static int64_t neighbours(const gview *g, uint64_t i, uint32_t *v_out, scratch_t *s) {
    if (!g->stream) {
        uint64_t d = g->row[i + 1] - g->row[i];
        if (grow(&s->nbr, &s->nbr_cap, d + 1)) return -9;
        memcpy(s->nbr, g->col + g->row[i], d * sizeof(uint32_t));
        *v_out = (uint32_t)i;
        return (int64_t)d;
    }
    uint32_t v = g->vertices[i];
    *v_out = v;
    if (i >= g->n_c) {
        uint64_t j = i - g->n_c;
        uint64_t d = g->soffs[j + 1] - g->soffs[j];
        if (grow(&s->nbr, &s->nbr_cap, d + 1)) return -9;
        memcpy(s->nbr, g->snbrs + g->soffs[j], d * sizeof(uint32_t));
        return (int64_t)d;
    }
    for (;;) {
        int64_t d = orc_decode_record(g->stream, g->nbits, g->offs[i], v, s->nbr, s->nbr_cap, NULL);
        if (d < 0) return d;
        if ((uint64_t)d <= s->nbr_cap) return d;
        if (grow(&s->nbr, &s->nbr_cap, (uint64_t)d)) return -9;
    }
}

static int64_t step_generic(const gview *g, const orc_lp_cfg *cfg, const uint8_t *seeds,
                            const uint64_t *list, uint64_t nlist,
                            const uint32_t *in, uint32_t *out) {
    if (cfg->K == 0) return -7;
    if (cfg->cms_thresh && (cfg->cms_depth == 0 || cfg->cms_depth > 16 || cfg->cms_width == 0 ||
                            (cfg->cms_width & (cfg->cms_width - 1)))) return -8;
    uint64_t ab[32] = {0};
    orc_cms_params(cfg->cms_seed, cfg->cms_depth, ab);
    int64_t changed = 0;
    int err = 0;
    uint64_t total = list ? nlist : g->n;
    int nt = cfg->nthreads > 0 ? cfg->nthreads : orc_max_threads();
    (void)nt;
#pragma omp parallel num_threads(nt) reduction(+ : changed)
    {
        scratch_t s = {0};
        if (cfg->cms_thresh) s.sketch = (uint32_t *)malloc((size_t)cfg->cms_depth * cfg->cms_width * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 64)
        for (uint64_t q = 0; q < total; q++) {
            if (err) continue;
            uint64_t i = list ? list[q] : q;
            uint32_t v;
            int64_t d = neighbours(g, i, &v, &s);
            if (d < 0) { err = (int)d; continue; }
            for (uint64_t j = 0; j < (uint64_t)d; j++)
                if (s.nbr[j] >= g->n) { err = -4; break; }
            if (err) continue;
            const uint64_t K = cfg->K;
            if (grow(&s.lab, &s.lab_cap, K * (uint64_t)d + 1) || grow(&s.tmp, &s.tmp_cap, (uint64_t)d + 1)) {
                err = -9;
                continue;
            }
            /* a neighbour's K labels are one row: gathered once, vector-major */
            for (uint64_t j = 0; j < (uint64_t)d; j++) {
                const uint32_t *rowp = in + (uint64_t)s.nbr[j] * K;
                for (uint64_t k = 0; k < K; k++) s.lab[k * (uint64_t)d + j] = rowp[k];
            }
            for (uint32_t k = 0; k < cfg->K; k++) {
                uint64_t idx = (uint64_t)v * cfg->K + k;
                uint32_t old = in[idx], nl;
                if (cfg->semi && seeds && seeds[idx]) nl = old;
                else nl = mode_of(cfg, ab, &s, s.lab + (uint64_t)k * d, (uint64_t)d, old);
                out[idx] = nl;
                changed += nl != old;
            }
        }
        free(s.nbr); free(s.lab); free(s.tmp); free(s.sketch);
    }
    return err ? err : changed;
}

int64_t orc_step_csr(uint64_t n, const uint64_t *row, const uint32_t *col,
                     const orc_lp_cfg *cfg, const uint8_t *seeds,
                     const uint32_t *in, uint32_t *out) {
    gview g = {0};
    g.n = n; g.row = row; g.col = col;
    return step_generic(&g, cfg, seeds, NULL, 0, in, out);
}

int64_t orc_step_cgr(uint64_t n, const uint8_t *stream, uint64_t nbits,
                     uint64_t n_c, const uint64_t *offs, const uint32_t *vertices,
                     const uint64_t *small_offs, const uint32_t *small_nbrs,
                     const orc_lp_cfg *cfg, const uint8_t *seeds,
                     const uint32_t *in, uint32_t *out) {
    gview g = {n, NULL, NULL, stream, nbits, n_c, offs, vertices, small_offs, small_nbrs};
    return step_generic(&g, cfg, seeds, NULL, 0, in, out);
}

int64_t orc_step_cgr_subset(uint64_t n, const uint8_t *stream, uint64_t nbits,
                            uint64_t n_c, const uint64_t *offs, const uint32_t *vertices,
                            const uint64_t *small_offs, const uint32_t *small_nbrs,
                            const orc_lp_cfg *cfg, const uint8_t *seeds,
                            const uint64_t *pos, uint64_t np,
                            const uint32_t *in, uint32_t *out) {
    gview g = {n, NULL, NULL, stream, nbits, n_c, offs, vertices, small_offs, small_nbrs};
    for (uint64_t q = 0; q < np; q++)
        if (pos[q] >= n) return -10;
    return step_generic(&g, cfg, seeds, pos, np, in, out);
}

int64_t orc_lp_run(uint64_t n, const uint64_t *row, const uint32_t *col,
                   const uint8_t *stream, uint64_t nbits, uint64_t n_c,
                   const uint64_t *offs, const uint32_t *vertices,
                   const uint64_t *small_offs, const uint32_t *small_nbrs,
                   const orc_lp_cfg *cfg, uint32_t *labels,
                   int32_t max_iters, int32_t until_converged, uint64_t *changed) {
    uint64_t tot = n * (uint64_t)cfg->K;
    uint32_t *tmp = (uint32_t *)malloc((tot ? tot : 1) * sizeof(uint32_t));
    uint8_t *seeds = NULL;
    if (!tmp) return -9;
    if (cfg->semi) {
        seeds = (uint8_t *)malloc(tot ? tot : 1);
        if (!seeds) { free(tmp); return -9; }
        for (uint64_t i = 0; i < tot; i++) seeds[i] = labels[i] != ORC_UNLABELLED;
    }
    gview g = {n, row, col, stream, nbits, n_c, offs, vertices, small_offs, small_nbrs};
    int64_t it = 0;
    for (; it < max_iters; it++) {
        int64_t c = step_generic(&g, cfg, seeds, NULL, 0, labels, tmp);
        if (c < 0) { free(tmp); free(seeds); return c; }
        memcpy(labels, tmp, tot * sizeof(uint32_t));
        if (changed) changed[it] = (uint64_t)c;
        if (until_converged && c == 0) { it++; break; }
    }
    free(tmp); free(seeds);
    return it;
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
