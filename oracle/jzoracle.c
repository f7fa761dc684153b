/*
 * jzoracle.c -- CPU restatement of jztree's tree hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (jztree_amd/) never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates (paths relative to
 * the jztree repository).  Where this build deliberately differs from the reference (kNN tie
 * order, z-order quirk for exactly equal coordinates) the difference is stated at the function.
 *
 * Build: oracle/Makefile  (gcc -O2 -fopenmp -ffp-contract=off -shared -fPIC)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------ */
/*                                    bit helpers                                             */
/* ------------------------------------------------------------------------------------------ */

static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static inline double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------------------------ */
/*                     z-order level math  (src/jztree_cuda/common/math.cuh)                 */
/* ------------------------------------------------------------------------------------------ */

/* math.cuh:241-269 msb_xor_float, literal (including the equal-bits quirk: clz(0)=32). */
int32_t orc_msb_xor_f32(float a, float b) {
    if (isnan(a) || isnan(b) || (signbit(a) != signbit(b))) return 128;
    uint32_t ab = f2u(fabsf(a)), bb = f2u(fabsf(b));
    int32_t ae = (int32_t)(ab >> 23) - 127, be = (int32_t)(bb >> 23) - 127;
    if (ae == be) {
        ae = imax(ae, -126);
        return ae + (8 - clz32(ab ^ bb));
    }
    return imax(ae, be);
}

/* math.cuh:271-290 msb_xor_double, literal.  (The reference has no NaN test here.) */
int32_t orc_msb_xor_f64(double a, double b) {
    if (signbit(a) != signbit(b)) return 1024;
    uint64_t ab = d2u(fabs(a)), bb = d2u(fabs(b));
    int32_t ae = (int32_t)(ab >> 52) - 1023, be = (int32_t)(bb >> 52) - 1023;
    if (ae == be) {
        ae = imax(ae, -1022);
        return ae + (11 - clz64(ab ^ bb));
    }
    return imax(ae, be);
}

/* math.cuh:313-327 msb_diff_level: (msb+1)*dim - msb_dim, highest level wins, ties -> lowest dim */
int32_t orc_msb_diff_level_f32(const float* p1, const float* p2, int dim) {
    int msb_dim = 0, msb = orc_msb_xor_f32(p1[0], p2[0]);
    for (int i = 1; i < dim; i++) {
        int m = orc_msb_xor_f32(p1[i], p2[i]);
        if (m > msb) { msb = m; msb_dim = i; }
    }
    return (msb + 1) * dim - msb_dim;
}
int32_t orc_msb_diff_level_f64(const double* p1, const double* p2, int dim) {
    int msb_dim = 0, msb = orc_msb_xor_f64(p1[0], p2[0]);
    for (int i = 1; i < dim; i++) {
        int m = orc_msb_xor_f64(p1[i], p2[i]);
        if (m > msb) { msb = m; msb_dim = i; }
    }
    return (msb + 1) * dim - msb_dim;
}

/* math.cuh:365-402 z_pos_less, literal (NaN rows last; compare along the dim of highest level). */
int orc_z_pos_less_f32(const float* p1, const float* p2, int dim) {
    int n1 = 0, n2 = 0;
    for (int i = 0; i < dim; i++) { n1 |= isnan(p1[i]); n2 |= isnan(p2[i]); }
    if (n1) return 0;
    if (n2) return 1;
    int msb_dim = 0, msb = orc_msb_xor_f32(p1[0], p2[0]);
    for (int i = 1; i < dim; i++) {
        int m = orc_msb_xor_f32(p1[i], p2[i]);
        if (m > msb) { msb = m; msb_dim = i; }
    }
    return p1[msb_dim] < p2[msb_dim];
}
int orc_z_pos_less_f64(const double* p1, const double* p2, int dim) {
    int n1 = 0, n2 = 0;
    for (int i = 0; i < dim; i++) { n1 |= isnan(p1[i]); n2 |= isnan(p2[i]); }
    if (n1) return 0;
    if (n2) return 1;
    int msb_dim = 0, msb = orc_msb_xor_f64(p1[0], p2[0]);
    for (int i = 1; i < dim; i++) {
        int m = orc_msb_xor_f64(p1[i], p2[i]);
        if (m > msb) { msb = m; msb_dim = i; }
    }
    return p1[msb_dim] < p2[msb_dim];
}

/*
 * Exact Morton order used by this build (a strict total order once the index breaks ties).
 *
 * It is z_pos_less (math.cuh:381-402) with two reference quirks removed:
 *  - exactly equal coordinates have no differing bit (level -inf) instead of level e-24
 *    (math.cuh:263-264, clz(0)=32);
 *  - -0.0 is treated as +0.0 (the reference gives the pair level 128 and calls it neither less
 *    nor greater).
 * Per dim the "top digit" ranks {-inf, negative finite, non-negative finite, +inf}; a differing
 * top digit is level 129 (above every finite magnitude bit, like the reference's 128 for sign
 * or inf differences).  Returns -1, 0, +1.
 */
static inline int topdig_f32(float x) {
    if (isinf(x)) return x < 0 ? 0 : 3;
    return (x < 0.0f) ? 1 : 2;
}
static inline int topdig_f64(double x) {
    if (isinf(x)) return x < 0 ? 0 : 3;
    return (x < 0.0) ? 1 : 2;
}
static inline int exact_level_f32(float a, float b) {
    int ta = topdig_f32(a), tb = topdig_f32(b);
    if (ta != tb) return 129;
    uint32_t ab = f2u(fabsf(a)), bb = f2u(fabsf(b));
    if (ab == bb) return -100000;
    int32_t ae = (int32_t)(ab >> 23) - 127, be = (int32_t)(bb >> 23) - 127;
    if (ae == be) return imax(ae, -126) + (8 - clz32(ab ^ bb));
    return imax(ae, be);
}
static inline int exact_level_f64(double a, double b) {
    int ta = topdig_f64(a), tb = topdig_f64(b);
    if (ta != tb) return 1025;
    uint64_t ab = d2u(fabs(a)), bb = d2u(fabs(b));
    if (ab == bb) return -100000;
    int32_t ae = (int32_t)(ab >> 52) - 1023, be = (int32_t)(bb >> 52) - 1023;
    if (ae == be) return imax(ae, -1022) + (11 - clz64(ab ^ bb));
    return imax(ae, be);
}
int orc_z_exact_cmp_f32(const float* p1, const float* p2, int dim) {
    int n1 = 0, n2 = 0;
    for (int i = 0; i < dim; i++) { n1 |= isnan(p1[i]); n2 |= isnan(p2[i]); }
    if (n1 || n2) return n1 - n2;
    int msb_dim = 0, msb = exact_level_f32(p1[0], p2[0]);
    for (int i = 1; i < dim; i++) {
        int m = exact_level_f32(p1[i], p2[i]);
        if (m > msb) { msb = m; msb_dim = i; }
    }
    if (msb == -100000) return 0;
    float a = p1[msb_dim], b = p2[msb_dim];
    if (msb == 129) return topdig_f32(a) < topdig_f32(b) ? -1 : 1;
    return a < b ? -1 : 1;
}
int orc_z_exact_cmp_f64(const double* p1, const double* p2, int dim) {
    int n1 = 0, n2 = 0;
    for (int i = 0; i < dim; i++) { n1 |= isnan(p1[i]); n2 |= isnan(p2[i]); }
    if (n1 || n2) return n1 - n2;
    int msb_dim = 0, msb = exact_level_f64(p1[0], p2[0]);
    for (int i = 1; i < dim; i++) {
        int m = exact_level_f64(p1[i], p2[i]);
        if (m > msb) { msb = m; msb_dim = i; }
    }
    if (msb == -100000) return 0;
    double a = p1[msb_dim], b = p2[msb_dim];
    if (msb == 1025) return topdig_f64(a) < topdig_f64(b) ? -1 : 1;
    return a < b ? -1 : 1;
}

/* Stable z-order sort: (exact Morton order, original index).  Restates PosZorderSort
 * (sort.cuh:40-76: CUB merge sort of (pos, iota) with z_pos_less); the reference leaves the
 * order of comparator-equivalent points to CUB, this build fixes it to the input index. */
static const float* g_sort_pos32;
static const double* g_sort_pos64;
static int g_sort_dim;
static int cmp_idx_f32(const void* a, const void* b) {
    int64_t i = *(const int64_t*)a, j = *(const int64_t*)b;
    int c = orc_z_exact_cmp_f32(g_sort_pos32 + i * g_sort_dim, g_sort_pos32 + j * g_sort_dim, g_sort_dim);
    if (c) return c;
    return (i > j) - (i < j);
}
static int cmp_idx_f64(const void* a, const void* b) {
    int64_t i = *(const int64_t*)a, j = *(const int64_t*)b;
    int c = orc_z_exact_cmp_f64(g_sort_pos64 + i * g_sort_dim, g_sort_pos64 + j * g_sort_dim, g_sort_dim);
    if (c) return c;
    return (i > j) - (i < j);
}
void orc_zsort_f32(const float* pos, int64_t n, int dim, int64_t* idz_out) {
    for (int64_t i = 0; i < n; i++) idz_out[i] = i;
    g_sort_pos32 = pos; g_sort_dim = dim;
    qsort(idz_out, (size_t)n, sizeof(int64_t), cmp_idx_f32);
}
void orc_zsort_f64(const double* pos, int64_t n, int dim, int64_t* idz_out) {
    for (int64_t i = 0; i < n; i++) idz_out[i] = i;
    g_sort_pos64 = pos; g_sort_dim = dim;
    qsort(idz_out, (size_t)n, sizeof(int64_t), cmp_idx_f64);
}

/* ------------------------------------------------------------------------------------------ */
/*                                tree kernels (tree.cuh)                                     */
/* ------------------------------------------------------------------------------------------ */

/*
 * tree.cuh:13-98 FlagLeafBoundaries, one boundary at a time.  Boundary b separates particles
 * b-1 and b.  level[b] = msb_diff_level(posz[b-1], posz[b]) with lvl_bound at the ends.
 * lsize/rsize count the particles of the binary Morton cell on each side (runs of boundaries
 * with a smaller level, window scan_size), ptype handling as in the reference (the type of the
 * particle left of each scanned boundary, tree.cuh:67-75).
 */
void orc_flag_leaf_boundaries_f32(const float* posz, const uint8_t* ptype, const int32_t* lvl_bound,
                                  int32_t npart, int64_t size_part, int dim, int max_size,
                                  int scan_size, int num_types, int8_t* split_flags, int32_t* lvl) {
    int64_t nb = size_part + 1;
    /* level of boundary b (b in [0, size_part]) */
    int32_t* L = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb));
    uint8_t* T = (uint8_t*)malloc((size_t)nb);
    for (int64_t b = 0; b < nb; b++) {
        int64_t ipart = b - 1;  /* particle left of boundary b */
        if (ipart < 0) { L[b] = lvl_bound[0]; T[b] = 0; }
        else if (ipart + 1 >= npart) { L[b] = lvl_bound[1]; T[b] = 0; }
        else {
            L[b] = orc_msb_diff_level_f32(posz + ipart * dim, posz + (ipart + 1) * dim, dim);
            T[b] = (num_types > 1) ? ptype[ipart] : 0;
        }
    }
    for (int64_t b = 0; b < nb; b++) {
        int my = L[b];
        int is_split = 0;
        for (int type = 0; type < num_types; type++) {
            int lsize = 0, rsize = 0;
            for (int64_t i = b - scan_size; i < b; i++) {
                int li = (i < 0) ? lvl_bound[0] : L[i];
                int ti = (i < 0) ? 0 : T[i];
                lsize = (li < my ? lsize : 0) + (ti == type);
            }
            for (int64_t i = b + scan_size; i > b; i--) {
                int li, ti;
                if (i >= nb) { li = lvl_bound[1]; ti = 0; }
                else { li = L[i]; ti = T[i]; }
                /* the reference's LDS window only holds boundaries <= size_part; beyond the
                 * particle range it stores lbound_r (ipart+1 >= nump), identical to L[i] here */
                rsize = (li < my ? rsize : 0) + (ti == type);
            }
            is_split |= (lsize + rsize > max_size);
            is_split |= ((my > lvl_bound[0]) && (b - lsize <= 0));
            is_split |= ((my > lvl_bound[1]) && (b + rsize >= npart));
        }
        is_split &= (b <= npart);
        is_split |= (b == 0);
        is_split |= (b == npart);
        split_flags[b] = (int8_t)is_split;
        lvl[b] = my;
    }
    free(L); free(T);
}

/* tree.cuh:104-199 FindNodeBoundaries, literal (binary searches with msb_diff_level). */
void orc_find_node_boundaries_f32(const float* pos_in, const float* pos_boundary, int32_t nleaves,
                                  int64_t size_nodes, int dim, int lvl_max, int lvl_invalid,
                                  int32_t* nodes_levels, int32_t* nodes_lbound, int32_t* nodes_rbound) {
    int n = nleaves, nnodes = n + 1;
    for (int64_t idx = 0; idx < size_nodes; idx++) {
        if (idx >= nnodes) {
            nodes_levels[idx] = lvl_invalid; nodes_lbound[idx] = n; nodes_rbound[idx] = n;
            continue;
        }
        const float* p1 = (idx == 0) ? pos_boundary : pos_in + (idx - 1) * dim;
        const float* p2 = (idx == nnodes - 1) ? pos_boundary + dim : pos_in + idx * dim;
        int target = orc_msb_diff_level_f32(p1, p2, dim);
        int lbound, rbound;
        if (orc_msb_diff_level_f32(pos_in, p2, dim) <= target) {
            lbound = 0;
        } else {
            int ibefore = 0, iinside = (int)idx;
            while (ibefore + 1 < iinside) {
                int itest = (ibefore + iinside) / 2;
                int l = orc_msb_diff_level_f32(pos_in + (int64_t)itest * dim, p2, dim);
                if (l > target) ibefore = itest; else iinside = itest;
            }
            lbound = iinside;
        }
        if (orc_msb_diff_level_f32(p1, pos_in + (int64_t)(n - 1) * dim, dim) <= target) {
            rbound = n;
        } else {
            int iinside = (int)idx - 1, iafter = n - 1;
            while (iinside + 1 < iafter) {
                int itest = (iinside + iafter) / 2;
                int l = orc_msb_diff_level_f32(p1, pos_in + (int64_t)itest * dim, dim);
                if (l > target) iafter = itest; else iinside = itest;
            }
            rbound = iafter;
        }
        nodes_levels[idx] = target; nodes_lbound[idx] = lbound; nodes_rbound[idx] = rbound;
    }
}

/* math.cuh:524-541 lvl_vec: python floor division of the binary level over dims */
static void lvl_vec(int level, int dim, int* lvec) {
    int olvl = (level + 2000 * dim) / dim - 2000;
    int omod = level - olvl * dim;
    for (int i = 0; i < dim; i++) lvec[i] = olvl + (omod >= (dim - i));
}

/* math.cuh:404-437 round_float_pow2_cent, literal */
static float round_f32_pow2_cent(float x, int level) {
    int32_t xb = (int32_t)f2u(fabsf(x));
    int32_t xe = (xb >> 23) - 127;
    if (level >= 128) return 0.f;
    if (level > xe) return ldexpf(signbit(x) ? -0.5f : 0.5f, level);
    int32_t keep = imax(xe, -126) - level;
    if (keep >= 23) return x;
    uint32_t mask = 0xFFFFFFFFu << (23 - keep);
    uint32_t nb = ((uint32_t)xb & mask) | (1u << (22 - keep));
    return signbit(x) ? -u2f(nb) : u2f(nb);
}

/* tree.cuh:312-372 GetNodeGeometry (lower- and upper-extent modes), literal. */
void orc_get_node_geometry_f32(const float* pos, const int32_t* lbound, const int32_t* rbound,
                               int32_t nnodes, int64_t size_nodes, int64_t size_part, int dim,
                               int lvl_invalid, uint32_t mode_flags, int upper_extent,
                               int32_t* level, float* center, float* extent) {
    for (int64_t idx = 0; idx < size_nodes; idx++) {
        if (idx >= nnodes) {
            if (mode_flags & 1u) level[idx] = lvl_invalid;
            if (mode_flags & 2u) for (int d = 0; d < dim; d++) center[idx * dim + d] = u2f(0x7fc00000u);
            if (mode_flags & 4u) for (int d = 0; d < dim; d++) extent[idx * dim + d] = 0.f;
            continue;
        }
        int lb = lbound[idx], rb = rbound[idx];
        int64_t i0 = imin(imax(lb, 0), (int)size_part - 1), i1 = imin(imax(rb - 1, 0), (int)size_part - 1);
        const float* x0b = pos + i0 * dim;
        const float* x1a = pos + i1 * dim;
        int lvl;
        if (!upper_extent) {
            lvl = orc_msb_diff_level_f32(x0b, x1a, dim);
        } else {
            float x0a[16], x1b[16];
            int64_t ia = imin(imax(lb - 1, 0), (int)size_part - 1), ib = imin(imax(rb, 0), (int)size_part - 1);
            for (int d = 0; d < dim; d++) {
                x0a[d] = (lb - 1 < 0) ? -INFINITY : pos[ia * dim + d];
                x1b[d] = (rb >= size_part) ? INFINITY : pos[ib * dim + d];
            }
            int ll = orc_msb_diff_level_f32(x0a, x0b, dim);
            int lr = orc_msb_diff_level_f32(x1a, x1b, dim);
            lvl = imin(ll, lr) - 1;
        }
        int lv[16];
        lvl_vec(lvl, dim, lv);
        if (mode_flags & 1u) level[idx] = lvl;
        if (mode_flags & 2u) for (int d = 0; d < dim; d++) center[idx * dim + d] = round_f32_pow2_cent(x0b[d], lv[d]);
        if (mode_flags & 4u) for (int d = 0; d < dim; d++) extent[idx * dim + d] = ldexpf(1.0f, lv[d]);
    }
}

/* ------------------------------------------------------------------------------------------ */
/*                              canonical distance (math.cuh:471-522)                         */
/* ------------------------------------------------------------------------------------------ */

/* wrap_dx (math.cuh:471-488): into [-L/2, L/2) by one add/subtract */
static inline float wrap_f32(float dx, float boxsize) {
    if (boxsize > 0.f) {
        float bh = 0.5f * boxsize;
        dx = dx < -bh ? dx + boxsize : dx;
        dx = dx >= bh ? dx - boxsize : dx;
    }
    return dx;
}
/* distance_squared (math.cuh:511-522) under nvcc's default contraction: an FMA chain
 * r2 = fma(dz,dz, fma(dy,dy, dx*dx)).  This is the canonical FP32 r^2 of this build. */
static inline float r2_f32(const float* a, const float* b, int dim, float boxsize) {
    float dx = wrap_f32(a[0] - b[0], boxsize);
    float r2 = dx * dx;
    for (int i = 1; i < dim; i++) {
        float d = wrap_f32(a[i] - b[i], boxsize);
        r2 = fmaf(d, d, r2);
    }
    return r2;
}
float orc_r2_f32(const float* a, const float* b, int dim, float boxsize) { return r2_f32(a, b, dim, boxsize); }

/* ------------------------------------------------------------------------------------------ */
/*                                    exact kNN                                               */
/* ------------------------------------------------------------------------------------------ */

/* Ordering of this build (differs from knn.cuh:84-102, which inserts with strict '<' in visit
 * order): neighbours ascending by (canonical r2, reported id). */
static inline int pair_less(float r2a, int64_t ia, float r2b, int64_t ib) {
    return (r2a < r2b) || (r2a == r2b && ia < ib);
}

static void topk_insert(float* r2s, int64_t* ids, int k, float r2, int64_t id) {
    if (!pair_less(r2, id, r2s[k - 1], ids[k - 1])) return;
    int i = k - 1;
    while (i > 0 && pair_less(r2, id, r2s[i - 1], ids[i - 1])) { r2s[i] = r2s[i - 1]; ids[i] = ids[i - 1]; i--; }
    r2s[i] = r2; ids[i] = id;
}

/* Brute force O(Nq*Ns): for every query, the k smallest (r2, id).  Output r2 (not sqrt). */
void orc_knn_brute_f32(const float* xs, const int64_t* ids, int64_t ns, const float* xq, int64_t nq,
                       int dim, int k, float boxsize, float* r2_out, int64_t* id_out) {
    /* one query per scheduling unit: a few queries against many sources still use every thread */
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nq; q++) {
        float* r2s = r2_out + q * k;
        int64_t* ii = id_out + q * k;
        for (int j = 0; j < k; j++) { r2s[j] = INFINITY; ii[j] = INT64_MAX; }
        for (int64_t s = 0; s < ns; s++) {
            float r2 = r2_f32(xq + q * dim, xs + s * dim, dim, boxsize);
            topk_insert(r2s, ii, k, r2, ids ? ids[s] : s);
        }
    }
}

/* Uniform-grid cell list over the bounding box (or the periodic box). */
typedef struct {
    int dim, nc[3];
    double lo[3], h[3];
    int64_t* cell_start;   /* ncells+1 */
    int64_t* order;        /* points sorted by cell */
    int periodic;
} grid_t;

static int64_t cell_of(const grid_t* g, const float* x, int c[3]) {
    int64_t lin = 0;
    for (int d = 0; d < g->dim; d++) {
        int cd = (int)floor(((double)x[d] - g->lo[d]) / g->h[d]);
        if (cd < 0) cd = 0;
        if (cd >= g->nc[d]) cd = g->nc[d] - 1;
        c[d] = cd;
        lin = lin * g->nc[d] + cd;
    }
    return lin;
}

static void grid_build(grid_t* g, const float* xs, int64_t ns, int dim, double target_h, float boxsize) {
    g->dim = dim;
    g->periodic = boxsize > 0.f;
    for (int d = 0; d < dim; d++) {
        double lo = INFINITY, hi = -INFINITY;
        if (g->periodic) { lo = 0.0; hi = boxsize; }
        else for (int64_t i = 0; i < ns; i++) { double v = xs[i * dim + d]; lo = v < lo ? v : lo; hi = v > hi ? v : hi; }
        if (!(hi > lo)) hi = lo + 1.0;
        double span = hi - lo;
        int nc = (int)floor(span / target_h);
        if (nc < 1) nc = 1;
        if (nc > 1024) nc = 1024;
        g->nc[d] = nc; g->lo[d] = lo; g->h[d] = span / nc;
    }
    int64_t ncell = 1;
    for (int d = 0; d < dim; d++) ncell *= g->nc[d];
    g->cell_start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
    g->order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ns > 0 ? ns : 1));
    int64_t* cid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ns > 0 ? ns : 1));
    int c[3];
    for (int64_t i = 0; i < ns; i++) { cid[i] = cell_of(g, xs + i * dim, c); g->cell_start[cid[i] + 1]++; }
    for (int64_t i = 0; i < ncell; i++) g->cell_start[i + 1] += g->cell_start[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    memcpy(fill, g->cell_start, sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < ns; i++) g->order[fill[cid[i]]++] = i;
    free(fill); free(cid);
}
static void grid_free(grid_t* g) { free(g->cell_start); free(g->order); }

/*
 * Exact kNN by expanding Chebyshev shells of grid cells.  A shell s (cells at Chebyshev
 * distance s from the query's cell) is only skipped once the lower bound (s-1)*h_min of its
 * distance, squared and with a 1e-5 relative margin for FP32 rounding of r2, exceeds the
 * current k-th r2.  Results equal orc_knn_brute_f32.
 */
void orc_knn_grid_f32(const float* xs, const int64_t* ids, int64_t ns, const float* xq, int64_t nq,
                      int dim, int k, float boxsize, float* r2_out, int64_t* id_out) {
    grid_t g;
    double vol = 1.0;
    for (int d = 0; d < dim; d++) {
        double lo = INFINITY, hi = -INFINITY;
        if (boxsize > 0.f) { lo = 0; hi = boxsize; }
        else for (int64_t i = 0; i < ns; i++) { double v = xs[i * dim + d]; lo = v < lo ? v : lo; hi = v > hi ? v : hi; }
        vol *= (hi > lo ? hi - lo : 1.0);
    }
    double per = (double)(k > 4 ? k : 4) * 0.5;
    double h = pow(vol * per / (double)(ns > 0 ? ns : 1), 1.0 / dim);
    grid_build(&g, xs, ns, dim, h, boxsize);
    double hmin = g.h[0];
    for (int d = 1; d < dim; d++) hmin = g.h[d] < hmin ? g.h[d] : hmin;
    int maxnc = 0;
    for (int d = 0; d < dim; d++) maxnc = imax(maxnc, g.nc[d]);

    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < nq; q++) {
        float* r2s = r2_out + q * k;
        int64_t* ii = id_out + q * k;
        for (int j = 0; j < k; j++) { r2s[j] = INFINITY; ii[j] = INT64_MAX; }
        int c0[3] = {0, 0, 0};
        cell_of(&g, xq + q * dim, c0);
        for (int s = 0; s <= maxnc; s++) {
            if (s >= 1) {
                double lb = (double)(s - 1) * hmin;
                if (lb * lb > (double)r2s[k - 1] * (1.0 + 1e-5)) break;
            }
            /* enumerate cells with Chebyshev distance exactly s */
            int lo[3], hi[3];
            for (int d = 0; d < 3; d++) { lo[d] = 0; hi[d] = 0; }
            for (int d = 0; d < dim; d++) { lo[d] = c0[d] - s; hi[d] = c0[d] + s; }
            int a[3];
            for (a[0] = lo[0]; a[0] <= hi[0]; a[0]++)
            for (a[1] = lo[1]; a[1] <= hi[1]; a[1]++)
            for (a[2] = lo[2]; a[2] <= hi[2]; a[2]++) {
                int cheb = 0;
                for (int d = 0; d < dim; d++) { int dd = abs(a[d] - c0[d]); cheb = imax(cheb, dd); }
                if (cheb != s) continue;
                int64_t lin = 0, skip = 0;
                for (int d = 0; d < dim; d++) {
                    int cd = a[d];
                    if (g.periodic) {
                        if (2 * s + 1 > g.nc[d] && (a[d] - c0[d] > g.nc[d] / 2 || a[d] - c0[d] < -((g.nc[d] - 1) / 2))) { skip = 1; break; }
                        cd = ((cd % g.nc[d]) + g.nc[d]) % g.nc[d];
                    } else if (cd < 0 || cd >= g.nc[d]) { skip = 1; break; }
                    lin = lin * g.nc[d] + cd;
                }
                if (skip) continue;
                for (int64_t t = g.cell_start[lin]; t < g.cell_start[lin + 1]; t++) {
                    int64_t si = g.order[t];
                    float r2 = r2_f32(xq + q * dim, xs + si * dim, dim, boxsize);
                    topk_insert(r2s, ii, k, r2, ids ? ids[si] : si);
                }
            }
        }
    }
    grid_free(&g);
}

/*
 * Exact kNN with a k-d tree (for clustered data, where a uniform grid degenerates).  Nodes split
 * the widest bounding-box dim at the median (quickselect); leaves hold <= 16 points.  A subtree is
 * skipped once the float64 lower bound of its box distance, squared and with a 1e-5 relative
 * margin for FP32 rounding of r2, exceeds the current k-th r2.  Results equal orc_knn_brute_f32.
 */
typedef struct { float lo[3], hi[3]; int64_t start, end; int64_t left, right; } kdnode_t;
typedef struct { kdnode_t* nodes; int64_t nnodes, cap; int64_t* idx; const float* x; int dim; } kdtree_t;

static void kd_box(kdtree_t* t, int64_t s, int64_t e, float* lo, float* hi) {
    for (int d = 0; d < t->dim; d++) { lo[d] = INFINITY; hi[d] = -INFINITY; }
    for (int64_t i = s; i < e; i++)
        for (int d = 0; d < t->dim; d++) {
            float v = t->x[t->idx[i] * t->dim + d];
            lo[d] = v < lo[d] ? v : lo[d];
            hi[d] = v > hi[d] ? v : hi[d];
        }
}
static void kd_select(kdtree_t* t, int64_t s, int64_t e, int64_t kth, int d) {
    int64_t* a = t->idx;
    const float* x = t->x;
    int dim = t->dim;
    while (e - s > 1) {
        float piv = x[a[s + (e - s) / 2] * dim + d];
        int64_t i = s, j = e - 1;
        while (i <= j) {
            while (x[a[i] * dim + d] < piv) i++;
            while (x[a[j] * dim + d] > piv) j--;
            if (i <= j) { int64_t tmp = a[i]; a[i] = a[j]; a[j] = tmp; i++; j--; }
        }
        if (kth <= j) e = j + 1;
        else if (kth >= i) s = i;
        else return;
    }
}
static int64_t kd_alloc(kdtree_t* t) {
    int64_t id;
    #pragma omp critical(kdalloc)
    { id = t->nnodes++; }
    return id;
}
static void kd_build_rec(kdtree_t* t, int64_t node, int64_t s, int64_t e, int depth) {
    kdnode_t* nd = &t->nodes[node];
    kd_box(t, s, e, nd->lo, nd->hi);
    nd->start = s; nd->end = e; nd->left = nd->right = -1;
    if (e - s <= 16) return;
    int bd = 0;
    float bw = -1.f;
    for (int d = 0; d < t->dim; d++) { float w = nd->hi[d] - nd->lo[d]; if (w > bw) { bw = w; bd = d; } }
    if (!(bw > 0.f)) return;  /* all points identical: keep as a (large) leaf */
    int64_t mid = s + (e - s) / 2;
    kd_select(t, s, e, mid, bd);
    int64_t l = kd_alloc(t), r = kd_alloc(t);
    t->nodes[node].left = l; t->nodes[node].right = r;
    if (depth < 6) {
        #pragma omp task
        kd_build_rec(t, l, s, mid, depth + 1);
        #pragma omp task
        kd_build_rec(t, r, mid, e, depth + 1);
        #pragma omp taskwait
    } else {
        kd_build_rec(t, l, s, mid, depth + 1);
        kd_build_rec(t, r, mid, e, depth + 1);
    }
}
static void kd_build(kdtree_t* t, const float* x, int64_t n, int dim) {
    t->x = x; t->dim = dim;
    t->cap = 2 * (n / 8 + 2) + 16;
    t->nodes = (kdnode_t*)malloc(sizeof(kdnode_t) * (size_t)t->cap);
    t->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) t->idx[i] = i;
    t->nnodes = 1;
    #pragma omp parallel
    #pragma omp single
    kd_build_rec(t, 0, 0, n, 0);
}
static double kd_mind2(const kdnode_t* nd, const float* q, int dim, float boxsize) {
    double r = 0.0;
    for (int d = 0; d < dim; d++) {
        double v = q[d], lo = nd->lo[d], hi = nd->hi[d], dd = 0.0;
        if (v < lo) dd = lo - v; else if (v > hi) dd = v - hi;
        if (boxsize > 0.f && dd > 0.0) {
            double L = boxsize;
            double a = (v + L < lo) ? lo - (v + L) : ((v + L > hi) ? v + L - hi : 0.0);
            double b = (v - L < lo) ? lo - (v - L) : ((v - L > hi) ? v - L - hi : 0.0);
            if (a < dd) dd = a;
            if (b < dd) dd = b;
        }
        r += dd * dd;
    }
    return r;
}
static void kd_query(const kdtree_t* t, int64_t node, const float* q, const int64_t* ids, int k, float boxsize,
                     float* r2s, int64_t* ii) {
    const kdnode_t* nd = &t->nodes[node];
    if (kd_mind2(nd, q, t->dim, boxsize) > (double)r2s[k - 1] * (1.0 + 1e-5)) return;
    if (nd->left < 0) {
        for (int64_t a = nd->start; a < nd->end; a++) {
            int64_t si = t->idx[a];
            float r2 = r2_f32(q, t->x + si * t->dim, t->dim, boxsize);
            topk_insert(r2s, ii, k, r2, ids ? ids[si] : si);
        }
        return;
    }
    double dl = kd_mind2(&t->nodes[nd->left], q, t->dim, boxsize);
    double dr = kd_mind2(&t->nodes[nd->right], q, t->dim, boxsize);
    if (dl <= dr) {
        kd_query(t, nd->left, q, ids, k, boxsize, r2s, ii);
        kd_query(t, nd->right, q, ids, k, boxsize, r2s, ii);
    } else {
        kd_query(t, nd->right, q, ids, k, boxsize, r2s, ii);
        kd_query(t, nd->left, q, ids, k, boxsize, r2s, ii);
    }
}
void orc_knn_kdtree_f32(const float* xs, const int64_t* ids, int64_t ns, const float* xq, int64_t nq,
                        int dim, int k, float boxsize, float* r2_out, int64_t* id_out) {
    kdtree_t t;
    kd_build(&t, xs, ns, dim);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < nq; q++) {
        float* r2s = r2_out + q * k;
        int64_t* ii = id_out + q * k;
        for (int j = 0; j < k; j++) { r2s[j] = INFINITY; ii[j] = INT64_MAX; }
        if (ns > 0) kd_query(&t, 0, xq + q * dim, ids, k, boxsize, r2s, ii);
    }
    free(t.nodes); free(t.idx);
}

/* ------------------------------------------------------------------------------------------ */
/*                 exact kNN in any dimension 2-10, float or double positions                  */
/* ------------------------------------------------------------------------------------------ */

/* wrap_dx<double> (math.cuh:471-488): the half box is a float (0.5 * boxsize, exact for the float
 * boxsize), the difference is wrapped in double by one add/subtract. */
static inline double wrap_f64(double dx, float boxsize) {
    if (boxsize > 0.f) {
        float bh = 0.5f * boxsize;
        dx = dx < -bh ? dx + (double)boxsize : dx;
        dx = dx >= bh ? dx - (double)boxsize : dx;
    }
    return dx;
}
/* distance_squared<dim,double> (math.cuh:511-522) under nvcc's contraction: r2 = dx*dx, then
 * r2 = fma(d, d, r2) for the further dims, in double. */
static inline double r2_f64(const double* a, const double* b, int dim, float boxsize) {
    double dx = wrap_f64(a[0] - b[0], boxsize);
    double r2 = dx * dx;
    for (int i = 1; i < dim; i++) {
        double d = wrap_f64(a[i] - b[i], boxsize);
        r2 = fma(d, d, r2);
    }
    return r2;
}
double orc_r2_f64(const double* a, const double* b, int dim, float boxsize) { return r2_f64(a, b, dim, boxsize); }

/* canonical r2 of source s and query q for either position type, as a double (a float r2 converts
 * exactly, so the (r2, id) order is the float one) */
static inline double r2_any(const void* xs, int64_t s, const void* xq, int64_t q, int dim, float boxsize, int is64) {
    if (is64) return r2_f64((const double*)xq + q * dim, (const double*)xs + s * dim, dim, boxsize);
    return (double)r2_f32((const float*)xq + q * dim, (const float*)xs + s * dim, dim, boxsize);
}
static inline int pair_less_d(double r2a, int64_t ia, double r2b, int64_t ib) {
    return (r2a < r2b) || (r2a == r2b && ia < ib);
}
static void topk_insert_d(double* r2s, int64_t* ids, int k, double r2, int64_t id) {
    if (!pair_less_d(r2, id, r2s[k - 1], ids[k - 1])) return;
    int i = k - 1;
    while (i > 0 && pair_less_d(r2, id, r2s[i - 1], ids[i - 1])) { r2s[i] = r2s[i - 1]; ids[i] = ids[i - 1]; i--; }
    r2s[i] = r2; ids[i] = id;
}

/* Brute force (knn.cuh:203-306 restated without a tree): per query the k smallest (canonical r2,
 * id), r2 as double.  Any dim, is64 = double positions. */
void orc_knn_brute_nd(const void* xs, const int64_t* ids, int64_t ns, const void* xq, int64_t nq, int dim, int k,
                      float boxsize, int is64, double* r2_out, int64_t* id_out) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nq; q++) {
        double* r2s = r2_out + q * k;
        int64_t* ii = id_out + q * k;
        for (int j = 0; j < k; j++) { r2s[j] = INFINITY; ii[j] = INT64_MAX; }
        for (int64_t s = 0; s < ns; s++)
            topk_insert_d(r2s, ii, k, r2_any(xs, s, xq, q, dim, boxsize, is64), ids ? ids[s] : s);
    }
}

/* k-d tree for dims up to 10 over a double copy of the positions (exact for float positions, so
 * the splits are those of the float values): median splits of the widest box dim, leaves of <= 16
 * points; a subtree is skipped once its float64 box distance^2 exceeds the current k-th r2 by a
 * 1e-5 relative margin (far above the rounding of either canonical r2).  Leaf points are scored
 * with the canonical r2 of the position type.  Results equal orc_knn_brute_nd. */
#define KDN_MAXD 10
typedef struct { double lo[KDN_MAXD], hi[KDN_MAXD]; int64_t start, end, left, right; } kdn_node_t;
typedef struct { kdn_node_t* nodes; int64_t nnodes, cap; int64_t* idx; const double* x; int dim; } kdn_tree_t;

static void kdn_box(kdn_tree_t* t, int64_t s, int64_t e, double* lo, double* hi) {
    for (int d = 0; d < t->dim; d++) { lo[d] = INFINITY; hi[d] = -INFINITY; }
    for (int64_t i = s; i < e; i++)
        for (int d = 0; d < t->dim; d++) {
            double v = t->x[t->idx[i] * t->dim + d];
            lo[d] = v < lo[d] ? v : lo[d];
            hi[d] = v > hi[d] ? v : hi[d];
        }
}
static void kdn_select(kdn_tree_t* t, int64_t s, int64_t e, int64_t kth, int d) {
    int64_t* a = t->idx;
    const double* x = t->x;
    int dim = t->dim;
    while (e - s > 1) {
        double piv = x[a[s + (e - s) / 2] * dim + d];
        int64_t i = s, j = e - 1;
        while (i <= j) {
            while (x[a[i] * dim + d] < piv) i++;
            while (x[a[j] * dim + d] > piv) j--;
            if (i <= j) { int64_t tmp = a[i]; a[i] = a[j]; a[j] = tmp; i++; j--; }
        }
        if (kth <= j) e = j + 1;
        else if (kth >= i) s = i;
        else return;
    }
}
static int64_t kdn_alloc(kdn_tree_t* t) {
    int64_t id;
    #pragma omp critical(kdnalloc)
    { id = t->nnodes++; }
    return id;
}
static void kdn_build_rec(kdn_tree_t* t, int64_t node, int64_t s, int64_t e, int depth) {
    kdn_node_t* nd = &t->nodes[node];
    kdn_box(t, s, e, nd->lo, nd->hi);
    nd->start = s; nd->end = e; nd->left = nd->right = -1;
    if (e - s <= 16) return;
    int bd = 0;
    double bw = -1.0;
    for (int d = 0; d < t->dim; d++) { double w = nd->hi[d] - nd->lo[d]; if (w > bw) { bw = w; bd = d; } }
    if (!(bw > 0.0)) return;  /* all points identical: keep as a (large) leaf */
    int64_t mid = s + (e - s) / 2;
    kdn_select(t, s, e, mid, bd);
    int64_t l = kdn_alloc(t), r = kdn_alloc(t);
    t->nodes[node].left = l; t->nodes[node].right = r;
    if (depth < 6) {
        #pragma omp task
        kdn_build_rec(t, l, s, mid, depth + 1);
        #pragma omp task
        kdn_build_rec(t, r, mid, e, depth + 1);
        #pragma omp taskwait
    } else {
        kdn_build_rec(t, l, s, mid, depth + 1);
        kdn_build_rec(t, r, mid, e, depth + 1);
    }
}
static double kdn_mind2(const kdn_node_t* nd, const double* q, int dim, float boxsize) {
    double r = 0.0;
    for (int d = 0; d < dim; d++) {
        double v = q[d], lo = nd->lo[d], hi = nd->hi[d], dd = 0.0;
        if (v < lo) dd = lo - v; else if (v > hi) dd = v - hi;
        if (boxsize > 0.f && dd > 0.0) {
            double L = boxsize;
            double a = (v + L < lo) ? lo - (v + L) : ((v + L > hi) ? v + L - hi : 0.0);
            double b = (v - L < lo) ? lo - (v - L) : ((v - L > hi) ? v - L - hi : 0.0);
            if (a < dd) dd = a;
            if (b < dd) dd = b;
        }
        r += dd * dd;
    }
    return r;
}
typedef struct { const void* xs; const void* xq; const int64_t* ids; int dim, k, is64; float boxsize; } kdn_q_t;
static void kdn_query(const kdn_tree_t* t, int64_t node, const kdn_q_t* c, int64_t q, const double* qd,
                      double* r2s, int64_t* ii) {
    const kdn_node_t* nd = &t->nodes[node];
    if (kdn_mind2(nd, qd, t->dim, c->boxsize) > r2s[c->k - 1] * (1.0 + 1e-5)) return;
    if (nd->left < 0) {
        for (int64_t a = nd->start; a < nd->end; a++) {
            int64_t si = t->idx[a];
            topk_insert_d(r2s, ii, c->k, r2_any(c->xs, si, c->xq, q, t->dim, c->boxsize, c->is64),
                          c->ids ? c->ids[si] : si);
        }
        return;
    }
    double dl = kdn_mind2(&t->nodes[nd->left], qd, t->dim, c->boxsize);
    double dr = kdn_mind2(&t->nodes[nd->right], qd, t->dim, c->boxsize);
    if (dl <= dr) {
        kdn_query(t, nd->left, c, q, qd, r2s, ii);
        kdn_query(t, nd->right, c, q, qd, r2s, ii);
    } else {
        kdn_query(t, nd->right, c, q, qd, r2s, ii);
        kdn_query(t, nd->left, c, q, qd, r2s, ii);
    }
}
static double* to_f64(const void* x, int64_t n, int is64) {
    double* d = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) d[i] = is64 ? ((const double*)x)[i] : (double)((const float*)x)[i];
    return d;
}
/* returns 0, or -1 for a dim outside 1..10 */
int orc_knn_kdtree_nd(const void* xs, const int64_t* ids, int64_t ns, const void* xq, int64_t nq, int dim, int k,
                      float boxsize, int is64, double* r2_out, int64_t* id_out) {
    if (dim < 1 || dim > KDN_MAXD) return -1;
    kdn_tree_t t;
    double* xsd = to_f64(xs, ns * dim, is64);
    double* xqd = to_f64(xq, nq * dim, is64);
    t.x = xsd; t.dim = dim;
    t.cap = 2 * (ns / 8 + 2) + 16;
    t.nodes = (kdn_node_t*)malloc(sizeof(kdn_node_t) * (size_t)t.cap);
    t.idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ns > 0 ? ns : 1));
    for (int64_t i = 0; i < ns; i++) t.idx[i] = i;
    t.nnodes = 1;
    #pragma omp parallel
    #pragma omp single
    kdn_build_rec(&t, 0, 0, ns, 0);
    kdn_q_t c = {xs, xq, ids, dim, k, is64, boxsize};
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < nq; q++) {
        double* r2s = r2_out + q * k;
        int64_t* ii = id_out + q * k;
        for (int j = 0; j < k; j++) { r2s[j] = INFINITY; ii[j] = INT64_MAX; }
        if (ns > 0) kdn_query(&t, 0, &c, q, xqd + q * dim, r2s, ii);
    }
    free(t.nodes); free(t.idx); free(xsd); free(xqd);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/*                                  friends-of-friends                                        */
/* ------------------------------------------------------------------------------------------ */

static int64_t uf_find(int64_t* p, int64_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}
static void uf_union(int64_t* p, int64_t a, int64_t b) {
    a = uf_find(p, a); b = uf_find(p, b);
    if (a == b) return;
    if (a < b) p[b] = a; else p[a] = b;
}

/*
 * FoF groups of the predicate canonical r2(i,j) <= r2link (fof.cuh:334-343 restated on all
 * pairs, no tree).  Output label[i] = min over i's group of key[j] (key = input index gives the
 * canonical labels; key = z-index gives the reference's z-order labels, fof.cuh:22-36).
 * The cell list uses cells >= rlink in every dim, so neighbours lie in adjacent cells.
 */
void orc_fof_f32(const float* x, int64_t n, int dim, float r2link, float boxsize,
                 const int64_t* key, int64_t* label_out) {
    int64_t* par = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) par[i] = i;
    double rl = sqrt((double)r2link) * (1.0 + 1e-6);
    grid_t g;
    grid_build(&g, x, n, dim, rl, boxsize);
    /* make sure every cell width is >= rlink (grid_build floors nc, so h >= target) */
    int64_t ncell = 1;
    for (int d = 0; d < dim; d++) ncell *= g.nc[d];
    int nnb = 1;
    for (int d = 0; d < dim; d++) nnb *= 3;
    for (int64_t cl = 0; cl < ncell; cl++) {
        int c[3] = {0, 0, 0};
        int64_t rem = cl;
        for (int d = dim - 1; d >= 0; d--) { c[d] = (int)(rem % g.nc[d]); rem /= g.nc[d]; }
        int64_t seen[27];
        int nseen = 0;
        for (int o = 0; o < nnb; o++) {
            int off[3] = {0, 0, 0}, oo = o, skip = 0;
            int64_t lin = 0;
            for (int d = 0; d < dim; d++) { off[d] = oo % 3 - 1; oo /= 3; }
            for (int d = 0; d < dim; d++) {
                int cd = c[d] + off[d];
                if (g.periodic) cd = ((cd % g.nc[d]) + g.nc[d]) % g.nc[d];
                else if (cd < 0 || cd >= g.nc[d]) { skip = 1; break; }
                lin = lin * g.nc[d] + cd;
            }
            if (skip || lin < cl) continue;
            int dup = 0;
            for (int s = 0; s < nseen; s++) dup |= (seen[s] == lin);
            if (dup) continue;
            seen[nseen++] = lin;
            for (int64_t a = g.cell_start[cl]; a < g.cell_start[cl + 1]; a++) {
                int64_t i = g.order[a];
                int64_t bstart = (lin == cl) ? a + 1 : g.cell_start[lin];
                for (int64_t b = bstart; b < g.cell_start[lin + 1]; b++) {
                    int64_t j = g.order[b];
                    if (r2_f32(x + i * dim, x + j * dim, dim, boxsize) <= r2link) uf_union(par, i, j);
                }
            }
        }
    }
    grid_free(&g);
    int64_t* mn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) mn[i] = INT64_MAX;
    for (int64_t i = 0; i < n; i++) {
        int64_t r = uf_find(par, i);
        int64_t kk = key ? key[i] : i;
        if (kk < mn[r]) mn[r] = kk;
    }
    for (int64_t i = 0; i < n; i++) label_out[i] = mn[uf_find(par, i)];
    free(mn); free(par);
}

/*
 * FoF on fine cells, the same groups as orc_fof_f32 (predicate canonical r2 <= r2link, labels =
 * min key per group) at a cost that does not blow up in dense clumps.
 *
 * Cells have side h <= sqrt(r2link/dim) * (1 - 2^-12).  Two points of one cell are then at true
 * squared distance <= r2link * (1 - 2^-11), and the canonical FP32 r2 of the pair (at most a few
 * roundings of relative size 2^-24 above the true value) stays below float32 r2link: every cell
 * is one group from the start.  Two cells within reach (index distance <= ceil(rlink/h) in every
 * dim) belong to one group as soon as one cross pair satisfies the predicate, so a cell pair is
 * scanned only until its first link.  The scan of cell pairs runs in parallel and records links;
 * a sequential union-find over cells joins them.  Points must lie in [0, boxsize) when periodic;
 * otherwise (or when the cells would be too many to index) the plain cell list is used.
 */
typedef struct { int64_t key; int64_t val; } kv64_t;

static int cmp_kv64(const void* a, const void* b) {
    const kv64_t* x = (const kv64_t*)a; const kv64_t* y = (const kv64_t*)b;
    return (x->key > y->key) - (x->key < y->key);
}

static inline uint64_t mix64(uint64_t z) {  /* splitmix64 finaliser (hash of a cell id) */
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void orc_fof_cells_f32(const float* x, int64_t n, int dim, float r2link, float boxsize,
                       const int64_t* key, int64_t* label_out) {
    const int periodic = boxsize > 0.f;
    double rl = sqrt((double)r2link);
    double hmax = rl / sqrt((double)dim) * (1.0 - 1.0 / 4096.0);
    double lo[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    int64_t nc[3] = {1, 1, 1};
    int ok = n > 0 && hmax > 0.0 && dim <= 3;
    for (int d = 0; d < dim && ok; d++) {
        double a = INFINITY, b = -INFINITY;
        for (int64_t i = 0; i < n; i++) { double v = x[i * dim + d]; a = v < a ? v : a; b = v > b ? v : b; }
        if (!(a >= -1e300 && b <= 1e300)) { ok = 0; break; }         /* NaN / inf */
        if (periodic) {
            if (a < 0.0 || b >= (double)boxsize) { ok = 0; break; }
            nc[d] = (int64_t)ceil((double)boxsize / hmax);
            if (nc[d] < 1) nc[d] = 1;
            lo[d] = 0.0; h[d] = (double)boxsize / (double)nc[d];
        } else {
            nc[d] = (int64_t)floor((b - a) / hmax) + 1;
            lo[d] = a; h[d] = hmax;
        }
        if (nc[d] > (1 << 20)) ok = 0;
    }
    if (!ok) { orc_fof_f32(x, n, dim, r2link, boxsize, key, label_out); return; }
    int R = 1;
    for (int d = 0; d < dim; d++) { int r = (int)ceil(rl / h[d]); R = r > R ? r : R; }

    /* points sorted by cell */
    kv64_t* pc = (kv64_t*)malloc(sizeof(kv64_t) * (size_t)n);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t lin = 0;
        for (int d = dim - 1; d >= 0; d--) {
            int64_t c = (int64_t)floor(((double)x[i * dim + d] - lo[d]) / h[d]);
            c = c < 0 ? 0 : (c >= nc[d] ? nc[d] - 1 : c);
            lin = lin * nc[d] + c;
        }
        pc[i].key = lin; pc[i].val = i;
    }
    qsort(pc, (size_t)n, sizeof(kv64_t), cmp_kv64);
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) if (i == 0 || pc[i].key != pc[i - 1].key) m++;
    int64_t* cstart = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m + 1));
    int64_t* cid = (int64_t*)malloc(sizeof(int64_t) * (size_t)m);
    m = 0;
    for (int64_t i = 0; i < n; i++)
        if (i == 0 || pc[i].key != pc[i - 1].key) { cid[m] = pc[i].key; cstart[m++] = i; }
    cstart[m] = n;
    /* open-addressing hash: cell id -> cell index */
    int64_t hs = 1;
    while (hs < 2 * m) hs <<= 1;
    int64_t* ht = (int64_t*)malloc(sizeof(int64_t) * (size_t)hs);
    for (int64_t i = 0; i < hs; i++) ht[i] = -1;
    for (int64_t c = 0; c < m; c++) {
        uint64_t s = mix64((uint64_t)cid[c]) & (uint64_t)(hs - 1);
        while (ht[s] >= 0) s = (s + 1) & (uint64_t)(hs - 1);
        ht[s] = c;
    }
    /* half stencil: offsets lexicographically after 0 */
    int noff = 0, offs[343][3];
    int span = 2 * R + 1, tot = 1;
    for (int d = 0; d < dim; d++) tot *= span;
    for (int o = 0; o < tot && noff < 343; o++) {
        int v[3] = {0, 0, 0}, oo = o, first = 0;
        for (int d = 0; d < dim; d++) { v[d] = oo % span - R; oo /= span; }
        for (int d = dim - 1; d >= 0; d--) if (v[d] != 0) { first = v[d]; break; }
        if (first > 0) { offs[noff][0] = v[0]; offs[noff][1] = v[1]; offs[noff][2] = v[2]; noff++; }
    }
    if (R > 3) { orc_fof_f32(x, n, dim, r2link, boxsize, key, label_out);
                 free(pc); free(cstart); free(cid); free(ht); return; }

    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    int64_t** lk = (int64_t**)calloc((size_t)nth, sizeof(int64_t*));
    int64_t* lkn = (int64_t*)calloc((size_t)nth, sizeof(int64_t));
    int64_t* lkc = (int64_t*)calloc((size_t)nth, sizeof(int64_t));
    #pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        #pragma omp for schedule(dynamic, 256)
        for (int64_t a = 0; a < m; a++) {
            int64_t ca[3] = {0, 0, 0}, rem = cid[a];
            for (int d = 0; d < dim; d++) { ca[d] = rem % nc[d]; rem /= nc[d]; }
            int64_t seen[343];
            int nseen = 0;
            for (int o = 0; o < noff; o++) {
                int64_t lin = 0;
                int skip = 0;
                for (int d = dim - 1; d >= 0; d--) {
                    int64_t c = ca[d] + offs[o][d];
                    if (periodic) c = ((c % nc[d]) + nc[d]) % nc[d];
                    else if (c < 0 || c >= nc[d]) { skip = 1; break; }
                    lin = lin * nc[d] + c;
                }
                if (skip || lin == cid[a]) continue;
                int dup = 0;
                for (int s = 0; s < nseen; s++) dup |= seen[s] == lin;
                if (dup) continue;
                seen[nseen++] = lin;
                uint64_t s = mix64((uint64_t)lin) & (uint64_t)(hs - 1);
                int64_t b = -1;
                while (ht[s] >= 0) { if (cid[ht[s]] == lin) { b = ht[s]; break; } s = (s + 1) & (uint64_t)(hs - 1); }
                if (b < 0) continue;
                /* (with few cells per dim a wrapped pair can be met from both cells: a redundant
                 * link, harmless to the union-find) */
                int linked = 0;
                for (int64_t i = cstart[a]; i < cstart[a + 1] && !linked; i++) {
                    const float* xi = x + pc[i].val * dim;
                    for (int64_t j = cstart[b]; j < cstart[b + 1]; j++) {
                        if (r2_f32(xi, x + pc[j].val * dim, dim, boxsize) <= r2link) { linked = 1; break; }
                    }
                }
                if (linked) {
                    if (lkn[t] + 2 > lkc[t]) {
                        lkc[t] = lkc[t] ? 2 * lkc[t] : 1024;
                        lk[t] = (int64_t*)realloc(lk[t], sizeof(int64_t) * (size_t)lkc[t]);
                    }
                    lk[t][lkn[t]++] = a; lk[t][lkn[t]++] = b;
                }
            }
        }
    }
    int64_t* par = (int64_t*)malloc(sizeof(int64_t) * (size_t)m);
    for (int64_t c = 0; c < m; c++) par[c] = c;
    for (int t = 0; t < nth; t++) {
        for (int64_t i = 0; i < lkn[t]; i += 2) uf_union(par, lk[t][i], lk[t][i + 1]);
        free(lk[t]);
    }
    free(lk); free(lkn); free(lkc);
    int64_t* mn = (int64_t*)malloc(sizeof(int64_t) * (size_t)m);
    for (int64_t c = 0; c < m; c++) mn[c] = INT64_MAX;
    for (int64_t c = 0; c < m; c++) {
        int64_t r = uf_find(par, c);
        for (int64_t i = cstart[c]; i < cstart[c + 1]; i++) {
            int64_t kk = key ? key[pc[i].val] : pc[i].val;
            if (kk < mn[r]) mn[r] = kk;
        }
    }
    for (int64_t c = 0; c < m; c++) {
        int64_t l = mn[uf_find(par, c)];
        for (int64_t i = cstart[c]; i < cstart[c + 1]; i++) label_out[pc[i].val] = l;
    }
    free(mn); free(par); free(pc); free(cstart); free(cid); free(ht);
}

/* brute force FoF (tiny n only): same output convention as orc_fof_f32 */
void orc_fof_brute_f32(const float* x, int64_t n, int dim, float r2link, float boxsize,
                       const int64_t* key, int64_t* label_out) {
    int64_t* par = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) par[i] = i;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++)
            if (r2_f32(x + i * dim, x + j * dim, dim, boxsize) <= r2link) uf_union(par, i, j);
    int64_t* mn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) mn[i] = INT64_MAX;
    for (int64_t i = 0; i < n; i++) {
        int64_t r = uf_find(par, i);
        int64_t kk = key ? key[i] : i;
        if (kk < mn[r]) mn[r] = kk;
    }
    for (int64_t i = 0; i < n; i++) label_out[i] = mn[uf_find(par, i)];
    free(mn); free(par);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* CenterOfMass (tree.cuh:378-429): per node i, the points [isplit[i], isplit[i+1]) in order,
 * sums of (m * x_d, m) with Kahan compensation when kahan != 0 -- y = fma(m, x_d, -c) for a
 * coordinate (one rounding, as a contracting compiler forms the reference's m * x followed by
 * "add - c"), y = m - c for the mass, t = sum + y, c = (t - sum) - y, sum = t -- or plain
 * sums (sum = fma(m, x_d, sum), sum + m) without.  out row = ((1 / M) * sum_d, M); NaN rows for
 * empty ranges and M <= 0.  -ffp-contract=off keeps every other operation as written. */
#define ORC_COM(T, FMA, NAME)                                                                         \
void NAME(const int32_t* isplit, const T* pos, const T* mass, int64_t nnodes, int dim, int kahan,     \
          T* out) {                                                                                   \
    for (int64_t i = 0; i < nnodes; i++) {                                                            \
        T sum[11], c[11];                                                                             \
        for (int d = 0; d <= dim; d++) sum[d] = c[d] = (T)0;                                          \
        const int32_t s0 = isplit[i], s1 = isplit[i + 1];                                             \
        for (int32_t p = s0; p < s1; p++) {                                                           \
            const T m = mass[p];                                                                      \
            for (int d = 0; d <= dim; d++) {                                                          \
                if (kahan) {                                                                          \
                    const T y = d < dim ? FMA(m, pos[(int64_t)p * dim + d], -c[d]) : m - c[d];        \
                    const T t = sum[d] + y;                                                           \
                    c[d] = (t - sum[d]) - y;                                                          \
                    sum[d] = t;                                                                       \
                } else {                                                                              \
                    sum[d] = d < dim ? FMA(m, pos[(int64_t)p * dim + d], sum[d]) : sum[d] + m;         \
                }                                                                                     \
            }                                                                                         \
        }                                                                                             \
        T* o = out + i * (dim + 1);                                                                   \
        if (s0 < s1 && sum[dim] > (T)0) {                                                             \
            const T inv = (T)1 / sum[dim];                                                            \
            for (int d = 0; d < dim; d++) o[d] = inv * sum[d];                                        \
            o[dim] = sum[dim];                                                                        \
        } else {                                                                                      \
            for (int d = 0; d <= dim; d++) o[d] = (T)NAN;                                             \
        }                                                                                             \
    }                                                                                                 \
}
ORC_COM(float, fmaf, orc_center_of_mass_f32)
ORC_COM(double, fma, orc_center_of_mass_f64)

/* Integer positions (PosZorderSort's int32_t / int64_t instances): z_pos_less (math.cuh:381-402)
 * with msb_xor (math.cuh:293-307) = index of the highest bit of a XOR b; the dimension with the
 * highest such bit decides (the first of equal ones) by signed comparison; equal points keep their
 * input order. */
static const int32_t* g_sort_pos_i32;
static const int64_t* g_sort_pos_i64;
static int zcmp_i64(const int64_t* a, const int64_t* b, int dim) {
    int md = -1, m = -1;
    for (int d = 0; d < dim; d++) {
        uint64_t x = (uint64_t)a[d] ^ (uint64_t)b[d];
        int l = x ? 63 - __builtin_clzll(x) : -1;
        if (l > m) { m = l; md = d; }
    }
    if (md < 0) return 0;
    return a[md] < b[md] ? -1 : 1;
}
static int zcmp_i32(const int32_t* a, const int32_t* b, int dim) {
    int md = -1, m = -1;
    for (int d = 0; d < dim; d++) {
        uint32_t x = (uint32_t)a[d] ^ (uint32_t)b[d];
        int l = x ? 31 - __builtin_clz(x) : -1;
        if (l > m) { m = l; md = d; }
    }
    if (md < 0) return 0;
    return a[md] < b[md] ? -1 : 1;
}
static int cmp_idx_i32(const void* a, const void* b) {
    int64_t i = *(const int64_t*)a, j = *(const int64_t*)b;
    int c = zcmp_i32(g_sort_pos_i32 + i * g_sort_dim, g_sort_pos_i32 + j * g_sort_dim, g_sort_dim);
    if (c) return c;
    return (i > j) - (i < j);
}
static int cmp_idx_i64(const void* a, const void* b) {
    int64_t i = *(const int64_t*)a, j = *(const int64_t*)b;
    int c = zcmp_i64(g_sort_pos_i64 + i * g_sort_dim, g_sort_pos_i64 + j * g_sort_dim, g_sort_dim);
    if (c) return c;
    return (i > j) - (i < j);
}
void orc_zsort_i32(const int32_t* pos, int64_t n, int dim, int64_t* idz_out) {
    for (int64_t i = 0; i < n; i++) idz_out[i] = i;
    g_sort_pos_i32 = pos; g_sort_dim = dim;
    qsort(idz_out, (size_t)n, sizeof(int64_t), cmp_idx_i32);
}
void orc_zsort_i64(const int64_t* pos, int64_t n, int dim, int64_t* idz_out) {
    for (int64_t i = 0; i < n; i++) idz_out[i] = i;
    g_sort_pos_i64 = pos; g_sort_dim = dim;
    qsort(idz_out, (size_t)n, sizeof(int64_t), cmp_idx_i64);
}
