/*
 * mwu_oracle.c -- CPU restatement of hpdex's Mann-Whitney U backend.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls
 * this file: it is the checker the GPU path is compared against (tests/),
 * the one-call check in __graft_entry__.smoke(), and the timed CPU baseline
 * in bench.py.  It follows the reference algorithm step by step:
 *
 *   gather by group   src/cpp/mannwhitneyu.cpp:913-926 (is_valid_value :611-615)
 *   sort ref, targets src/cpp/mannwhitneyu.cpp:880-886, 928-932 (VQSortStatic
 *                     replaced by an ascending quicksort; the merge only needs
 *                     ascending order under '<', so any correct sort gives the
 *                     same runs and midranks)
 *   zero modes        src/cpp/mannwhitneyu.cpp:935-1039
 *   merges            src/cpp/mannwhitneyu.cpp:618-659 (segment), :662-725 (mix),
 *                     :728-788 (min/max), :791-840 (none)
 *   asymptotic p      src/cpp/mannwhitneyu.cpp:20-96 (precise erfc path)
 *   exact p           structure of src/cpp/mannwhitneyu.cpp:98-188, :243-282
 *                     with the CORRECT Mann-Whitney null (the reference's
 *                     recurrence at :134-158 counts sums of iid uniforms, not
 *                     rank sums); counts follow the recurrence of
 *                     src/hpdex/backend/pykernel_.py:150-199, in FP64.
 *   group_mean        src/cpp/mannwhitneyu.cpp:1118-1192 (+ Kahan :1195-1272)
 *   errors            src/cpp/mannwhitneyu.cpp:938-1045, :1088-1090
 *
 * Deterministic error choice: the reference keeps the first error any thread
 * hits; with threads=1 that is the lowest column, reference group before
 * target groups, targets in order.  This restatement always reports that one.
 *
 * Pinning (tests/test_oracle_*.py): U1 and the tie term against scipy
 * reductions and brute-force midranks; asymptotic p against scipy's
 * formula; exact counts against brute-force enumeration and scipy's exact
 * null; group_mean against numpy.  fast_norm=True is not restated: the
 * reference's SIMD erfc is lane-dependent (parity unpinned), and the precise
 * path is used for both flag values.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* dtype codes: same numbering as include/hpdex_mi355x.h */
enum {
    DT_I8 = 0, DT_U8 = 1, DT_I16 = 2, DT_U16 = 3, DT_I32 = 4,
    DT_U32 = 5, DT_I64 = 6, DT_U64 = 7, DT_F32 = 8, DT_F64 = 9
};
enum { ZH_NONE = 0, ZH_MIN = 1, ZH_MAX = 2, ZH_MIX = 3 };
enum { ALT_LESS = 0, ALT_GREATER = 1, ALT_TWO_SIDED = 2 };
enum { METHOD_EXACT = 1, METHOD_ASYMPTOTIC = 2 };

static double load_as_double(int dtype, const void* data, size_t p) {
    switch (dtype) {
    case DT_I8:  return (double)((const int8_t*)data)[p];
    case DT_U8:  return (double)((const uint8_t*)data)[p];
    case DT_I16: return (double)((const int16_t*)data)[p];
    case DT_U16: return (double)((const uint16_t*)data)[p];
    case DT_I32: return (double)((const int32_t*)data)[p];
    case DT_U32: return (double)((const uint32_t*)data)[p];
    case DT_I64: return (double)((const int64_t*)data)[p];
    case DT_U64: return (double)((const uint64_t*)data)[p];
    case DT_F32: return (double)((const float*)data)[p];
    default:     return ((const double*)data)[p];
    }
}

/* is_valid_value (mannwhitneyu.cpp:611-615): floats must be finite. */
static int is_valid(int dtype, double v) {
    if (dtype == DT_F32 || dtype == DT_F64) return isfinite(v);
    return 1;
}

/* ---------------- ascending sort of doubles (no NaN present) ---------------- */
static void insertion_sort(double* a, size_t n) {
    for (size_t i = 1; i < n; ++i) {
        double v = a[i];
        size_t j = i;
        while (j > 0 && v < a[j - 1]) { a[j] = a[j - 1]; --j; }
        a[j] = v;
    }
}

static void sort_doubles(double* a, size_t n) {
    while (n > 24) {
        double x = a[0], y = a[n / 2], z = a[n - 1];
        double piv = (x < y) ? ((y < z) ? y : ((x < z) ? z : x))
                             : ((x < z) ? x : ((y < z) ? z : y));
        size_t i = 0, j = n - 1;
        for (;;) {
            while (a[i] < piv) ++i;
            while (piv < a[j]) --j;
            if (i >= j) break;
            double t = a[i]; a[i] = a[j]; a[j] = t;
            ++i; --j;
        }
        /* recurse on the smaller part, loop on the larger */
        size_t left = j + 1;
        if (left < n - left) { sort_doubles(a, left); a += left; n -= left; }
        else { sort_doubles(a + left, n - left); n = left; }
    }
    insertion_sort(a, n);
}

/* ---------------- merges (mannwhitneyu.cpp:618-840) ---------------- */
/* merge_and_rank_segment (:618-659) */
static size_t merge_segment(const double* a, size_t na, const double* b, size_t nb,
                            size_t rank, double* R1, double* tie_sum) {
    size_t i = 0, j = 0;
    while (i < na || j < nb) {
        double v;
        int take_a = 0, take_b = 0;
        if (i < na && j < nb) {
            if (a[i] < b[j]) { v = a[i]; take_a = 1; }
            else if (b[j] < a[i]) { v = b[j]; take_b = 1; }
            else { v = a[i]; take_a = take_b = 1; }
        } else if (i < na) { v = a[i]; take_a = 1; }
        else { v = b[j]; take_b = 1; }
        size_t eq_a = 0, eq_b = 0;
        if (take_a) while (i + eq_a < na && !(a[i + eq_a] < v) && !(v < a[i + eq_a])) ++eq_a;
        if (take_b) while (j + eq_b < nb && !(b[j + eq_b] < v) && !(v < b[j + eq_b])) ++eq_b;
        size_t t = eq_a + eq_b;
        double start = (double)rank, end = (double)(rank + t - 1);
        double avg = 0.5 * (start + end);
        *R1 += (double)eq_a * avg;
        if (t > 1) { double tt = (double)t; *tie_sum += (tt * tt * tt - tt); }
        rank += t;
        i += eq_a; j += eq_b;
    }
    return rank;
}

/* neg / zero / pos split used by :672-687 and :739-754 */
static void split3(const double* a, size_t n, size_t* neg_end, size_t* zero_end) {
    size_t e = 0;
    while (e < n && a[e] < 0.0) ++e;
    *neg_end = e;
    while (e < n && !(0.0 < a[e]) && !(a[e] < 0.0)) ++e;
    *zero_end = e;
}

static void emit_zero_block(size_t* rank, size_t a_zero, size_t t_zero,
                            double* R1, double* tie_sum) {
    if (t_zero > 0) {
        double start = (double)*rank, end = (double)(*rank + t_zero - 1);
        double avg = 0.5 * (start + end);
        *R1 += (double)a_zero * avg;
        if (t_zero > 1) { double tz = (double)t_zero; *tie_sum += (tz * tz * tz - tz); }
        *rank += t_zero;
    }
}

/* min/max (:728-788) and mix (:662-725): zero block = explicit + implicit zeros */
static void merge_zero_modes(const double* a, size_t na, const double* b, size_t nb,
                             size_t a_zero_imp, size_t b_zero_imp, int mode,
                             double* R1, double* tie_sum) {
    size_t a_ne, a_ze, b_ne, b_ze;
    split3(a, na, &a_ne, &a_ze);
    split3(b, nb, &b_ne, &b_ze);
    size_t A_zero = (a_ze - a_ne) + a_zero_imp;
    size_t B_zero = (b_ze - b_ne) + b_zero_imp;
    size_t t_zero = A_zero + B_zero;
    size_t rank = 1;
    *R1 = 0.0; *tie_sum = 0.0;
    if (mode == ZH_MIN) emit_zero_block(&rank, A_zero, t_zero, R1, tie_sum);
    rank = merge_segment(a, a_ne, b, b_ne, rank, R1, tie_sum);
    if (mode == ZH_MIX) emit_zero_block(&rank, A_zero, t_zero, R1, tie_sum);
    rank = merge_segment(a + a_ze, na - a_ze, b + b_ze, nb - b_ze, rank, R1, tie_sum);
    if (mode == ZH_MAX) emit_zero_block(&rank, A_zero, t_zero, R1, tie_sum);
}

/* ---------------- asymptotic p (mannwhitneyu.cpp:20-96) ---------------- */
static void mu_sigma(double n1, double n2, double tie_sum, double* mu, double* sigma) {
    const double N = n1 + n2;
    const double nm = n1 * n2;
    *mu = 0.5 * nm;
    double tie_term = 0.0;
    const double denom = N * (N - 1.0);
    if (denom > 0.0) tie_term = tie_sum / denom;
    const double var = nm * ((N + 1.0) - tie_term) / 12.0;
    *sigma = (var > 0.0) ? sqrt(var) : 0.0;
}

static double norm_sf(double z) { return 0.5 * erfc(z * 0.70710678118654752440); }

double oracle_p_asymptotic(double U1, double n1, double n2, double tie, double cc, int alt) {
    double mu, sigma, p = 1.0;
    const double nm = n1 * n2;
    mu_sigma(n1, n2, tie, &mu, &sigma);
    if (alt == ALT_TWO_SIDED) {
        const double U2 = nm - U1;
        const double U = (U1 > U2 ? U1 : U2);
        if (sigma > 0.0) {
            const double z = (U - mu - cc) / sigma;
            p = 2.0 * norm_sf(z);
            if (p > 1.0) p = 1.0;
            if (p < 0.0) p = 0.0;
        }
        return p;
    }
    const double U = (alt == ALT_GREATER) ? U1 : (nm - U1);
    if (sigma > 0.0) {
        const double z = (U - mu - cc) / sigma;
        p = norm_sf(z);
    }
    return (p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p));
}

/* ---------------- exact null (correct Mann-Whitney distribution) ----------------
 * counts[k] = number of arrangements of n1 and n2 items whose U equals k,
 * k = 0..n1*n2.  Recurrence on the largest item (pykernel_.py:173-185):
 *   f_i[j][k] = f_i[j-1][k] + f_{i-1}[j][k-j],  f_i[0][0] = f_0[j][0] = 1,
 * with i running over the larger sample and j over the smaller one (the null
 * is symmetric in n1, n2).  Every entry is one FP64 addition of non-negative
 * terms, so the order of evaluation inside a sweep does not change any bit.
 * Returns 0 on success, -1 if the table would exceed max_cells doubles.   */
int oracle_exact_counts(size_t n1, size_t n2, double* counts, size_t max_cells) {
    const size_t s = n1 < n2 ? n1 : n2;
    const size_t l = n1 < n2 ? n2 : n1;
    const size_t U = n1 * n2;
    const size_t W = U + 1;
    if ((s + 1) * W > max_cells) return -1;
    double* fp = (double*)calloc((s + 1) * W, sizeof(double));
    double* fc = (double*)calloc((s + 1) * W, sizeof(double));
    if (!fp || !fc) { free(fp); free(fc); return -1; }
    for (size_t j = 0; j <= s; ++j) fp[j * W] = 1.0;
    for (size_t i = 1; i <= l; ++i) {
        const size_t cap = i * s;
        memset(fc, 0, (s + 1) * W * sizeof(double));
        fc[0] = 1.0;
        for (size_t j = 1; j <= s; ++j) {
            double* cur = fc + j * W;
            const double* below = fc + (j - 1) * W;
            const double* prev = fp + j * W;
            for (size_t k = 0; k <= cap; ++k)
                cur[k] = (k >= j) ? below[k] + prev[k - j] : below[k];
        }
        double* t = fp; fp = fc; fc = t;
    }
    memcpy(counts, fp + s * W, W * sizeof(double));
    free(fp); free(fc);
    return 0;
}

/* p_exact structure of mannwhitneyu.cpp:98-188 on the correct null. */
static double sf_from_counts(const double* c, size_t Umax, double Ustat) {
    if (Umax == 0) return 1.0;
    double Uc = Ustat < 0.0 ? 0.0 : (Ustat > (double)Umax ? (double)Umax : Ustat);
    const size_t u = (size_t)floor(Uc);
    double total = 0.0;
    for (size_t k = 0; k <= Umax; ++k) total += c[k];
    const size_t kc = Umax - u;
    const size_t small = u < kc ? u : kc;
    double cdf_small = 0.0;
    for (size_t k = 0; k <= small; ++k) cdf_small += c[k];
    const double pmf_small = c[small];
    if (u <= kc) return 1.0 - cdf_small / total + pmf_small / total;
    return cdf_small / total;
}

typedef struct { size_t n1, n2; double* counts; } exact_entry;
typedef struct { exact_entry* e; size_t n, cap; } exact_cache;

static const double* cache_get(exact_cache* cc, size_t n1, size_t n2) {
    for (size_t i = 0; i < cc->n; ++i)
        if (cc->e[i].n1 == n1 && cc->e[i].n2 == n2) return cc->e[i].counts;
    double* c = (double*)malloc((n1 * n2 + 1) * sizeof(double));
    if (!c) return NULL;
    if (oracle_exact_counts(n1, n2, c, (size_t)1 << 28) != 0) { free(c); return NULL; }
    if (cc->n == cc->cap) {
        cc->cap = cc->cap ? 2 * cc->cap : 16;
        cc->e = (exact_entry*)realloc(cc->e, cc->cap * sizeof(exact_entry));
    }
    cc->e[cc->n].n1 = n1; cc->e[cc->n].n2 = n2; cc->e[cc->n].counts = c;
    return cc->e[cc->n++].counts;
}

/* p_exact_parallel (:243-282) body for one pair */
double oracle_p_exact(double U1, size_t n1, size_t n2, const double* counts, int alt) {
    const double nm = (double)n1 * (double)n2;
    const size_t Umax = n1 * n2;
    double p = 1.0;
    if (alt == ALT_TWO_SIDED) {
        const double U2 = nm - U1;
        p = 2.0 * sf_from_counts(counts, Umax, U1 > U2 ? U1 : U2);
    } else if (alt == ALT_GREATER) {
        p = sf_from_counts(counts, Umax, U1);
    } else {
        p = sf_from_counts(counts, Umax, nm - U1);
    }
    if (p < 0.0) p = 0.0;
    if (p > 1.0) p = 1.0;
    return p;
}

#define OVERFLOW_TAG ((size_t)1 << 40)

static void set_err(char* err, size_t len, const char* msg) {
    if (err && len) { strncpy(err, msg, len - 1); err[len - 1] = 0; }
}

/* ---------------- main entry (mannWhitneyu_core, :843-1065) ----------------
 * Outputs (each C*n_targets doubles, gene-major idx = c*T + (g-1)); U1 and P
 * required, tie/n1/n2 optional (NULL to skip).  Returns 0 or 1 with err set.  */
int oracle_mwu(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
               size_t R, size_t C, size_t nnz, const int32_t* group_id, size_t n_targets,
               int tie_correction, int use_continuity, int zero_handling,
               int alternative, int method, int threads,
               double* U1_out, double* P_out, double* tie_out, double* n1_out, double* n2_out,
               char* err, size_t err_len) {
    (void)nnz;
    if (n_targets < 1) {
        set_err(err, err_len, "[mannwhitney] n_targets must be >= 1 (>=1 target groups)");
        return 1;
    }
    if (dtype < DT_I8 || dtype > DT_F64) { set_err(err, err_len, "Unsupported dtype for data"); return 1; }
    const size_t G = n_targets + 1;
    const size_t Np = C * n_targets;
    size_t* group_rows = (size_t*)calloc(G, sizeof(size_t));
    size_t* group_off = (size_t*)calloc(G + 1, sizeof(size_t));
    double* n1a = (double*)calloc(Np ? Np : 1, sizeof(double));
    double* n2a = (double*)calloc(Np ? Np : 1, sizeof(double));
    double* tia = (double*)calloc(Np ? Np : 1, sizeof(double));
    /* per-column first error code: 0 none, 1 ref, 2+g target g */
    size_t* col_err = (size_t*)calloc(C ? C : 1, sizeof(size_t));
    for (size_t r = 0; r < R; ++r) {
        int g = group_id[r];
        if (g >= 0 && (size_t)g < G) ++group_rows[g];
    }
    for (size_t g = 1; g <= G; ++g) group_off[g] = group_off[g - 1] + group_rows[g - 1];
    const size_t cap_total = group_off[G];
    const double tie_mask = tie_correction ? 1.0 : 0.0;
    for (size_t i = 0; i < Np; ++i) { U1_out[i] = 0.0; P_out[i] = 1.0; }
    if (threads <= 0) threads = 1;
#ifdef _OPENMP
    if (threads > omp_get_max_threads()) threads = omp_get_max_threads();
#endif
#pragma omp parallel num_threads(threads)
    {
        double* buf = (double*)malloc((cap_total ? cap_total : 1) * sizeof(double));
        size_t* cnt = (size_t*)malloc(G * sizeof(size_t));
#pragma omp for schedule(dynamic, 4)
        for (long long cc = 0; cc < (long long)C; ++cc) {
            memset(cnt, 0, G * sizeof(size_t));
            for (int64_t p = indptr[cc]; p < indptr[cc + 1]; ++p) {
                const int64_t r = indices[p];
                if (r < 0 || (size_t)r >= R) continue;
                const int g = group_id[r];
                if (g < 0 || (size_t)g >= G) continue;
                const double v = load_as_double(dtype, data, (size_t)p);
                if (!is_valid(dtype, v)) continue;
                if (cnt[g] >= group_rows[g]) {
                    /* duplicate row indices overflow the group's capacity:
                     * undefined in the reference, an error here */
                    if (!col_err[cc]) col_err[cc] = OVERFLOW_TAG + (size_t)g;
                    continue;
                }
                buf[group_off[g] + cnt[g]++] = v;
            }
            if (col_err[cc]) continue;
            for (size_t g = 0; g < G; ++g)
                if (cnt[g] > 1) sort_doubles(buf + group_off[g], cnt[g]);
            const double* a = buf;
            const size_t n1_total = (zero_handling == ZH_NONE) ? cnt[0] : group_rows[0];
            if (n1_total < 2) { col_err[cc] = 1; continue; }
            const double n1d = (double)n1_total;
            for (size_t g = 1; g < G; ++g) {
                const size_t n2_total = (zero_handling == ZH_NONE) ? cnt[g] : group_rows[g];
                if (n2_total < 2) { if (!col_err[cc]) col_err[cc] = 1 + g; continue; }
                double R1 = 0.0, tsum = 0.0;
                const double* b = buf + group_off[g];
                if (zero_handling == ZH_NONE) {
                    merge_segment(a, cnt[0], b, cnt[g], 1, &R1, &tsum);
                } else {
                    merge_zero_modes(a, cnt[0], b, cnt[g], group_rows[0] - cnt[0],
                                     group_rows[g] - cnt[g], zero_handling, &R1, &tsum);
                }
                const size_t idx = (size_t)cc * n_targets + (g - 1);
                U1_out[idx] = R1 - n1d * (n1d + 1.0) * 0.5;
                n1a[idx] = n1d;
                n2a[idx] = (double)n2_total;
                tia[idx] = tsum * tie_mask;
            }
        }
        free(buf); free(cnt);
    }
    int rc = 0;
    for (size_t c = 0; c < C; ++c) {
        if (col_err[c]) {
            char msg[256];
            if (col_err[c] >= OVERFLOW_TAG)
                snprintf(msg, sizeof msg,
                         "Non-canonical CSC: column %zu holds more entries for group %zu than the group has rows",
                         c, col_err[c] - OVERFLOW_TAG);
            else if (col_err[c] == 1) snprintf(msg, sizeof msg, "Sample too small for reference at column %zu", c);
            else snprintf(msg, sizeof msg, "Sample too small for group %zu at column %zu", col_err[c] - 1, c);
            set_err(err, err_len, msg);
            rc = 1;
            break;
        }
    }
    if (rc == 0) {
        const double ccv = use_continuity ? 0.5 : 0.0;
        if (method == METHOD_ASYMPTOTIC) {
#pragma omp parallel for num_threads(threads) schedule(static)
            for (long long i = 0; i < (long long)Np; ++i)
                P_out[i] = oracle_p_asymptotic(U1_out[i], n1a[i], n2a[i], tia[i], ccv, alternative);
        } else {
            exact_cache cache = {0};
            for (size_t i = 0; i < Np; ++i) {
                const double* c = cache_get(&cache, (size_t)n1a[i], (size_t)n2a[i]);
                if (!c) { set_err(err, err_len, "exact null table too large"); rc = 1; break; }
                P_out[i] = oracle_p_exact(U1_out[i], (size_t)n1a[i], (size_t)n2a[i], c, alternative);
            }
            for (size_t i = 0; i < cache.n; ++i) free(cache.e[i].counts);
            free(cache.e);
        }
        if (tie_out) memcpy(tie_out, tia, Np * sizeof(double));
        if (n1_out) memcpy(n1_out, n1a, Np * sizeof(double));
        if (n2_out) memcpy(n2_out, n2a, Np * sizeof(double));
    }
    free(group_rows); free(group_off); free(n1a); free(n2a); free(tia); free(col_err);
    return rc;
}

/* ---------------- group_mean (mannwhitneyu.cpp:1118-1272) ---------------- */
int oracle_group_mean(int dtype, const void* data, const int64_t* indices, const int64_t* indptr,
                      size_t R, size_t C, size_t nnz, const int32_t* group_id, size_t n_groups,
                      int include_zeros, int use_kahan, int threads, double* out,
                      char* err, size_t err_len) {
    (void)nnz;
    if (dtype < DT_I8 || dtype > DT_F64) { set_err(err, err_len, "Unsupported dtype for data"); return 1; }
    const size_t G = n_groups;
    if (G == 0 || C == 0) return 0;
    size_t* group_rows = (size_t*)calloc(G, sizeof(size_t));
    for (size_t r = 0; r < R; ++r) {
        int g = group_id[r];
        if (g >= 0 && (size_t)g < G) ++group_rows[g];
    }
    if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads)
    {
        double* sum = (double*)malloc(G * sizeof(double));
        double* comp = (double*)malloc(G * sizeof(double));
        size_t* vc = (size_t*)malloc(G * sizeof(size_t));
        size_t* ic = (size_t*)malloc(G * sizeof(size_t));
#pragma omp for schedule(dynamic, 4)
        for (long long cc = 0; cc < (long long)C; ++cc) {
            memset(sum, 0, G * sizeof(double));
            memset(comp, 0, G * sizeof(double));
            memset(vc, 0, G * sizeof(size_t));
            memset(ic, 0, G * sizeof(size_t));
            for (int64_t p = indptr[cc]; p < indptr[cc + 1]; ++p) {
                const int64_t r = indices[p];
                if (r < 0 || (size_t)r >= R) continue;
                const int g = group_id[r];
                if (g < 0 || (size_t)g >= G) continue;
                const double v = load_as_double(dtype, data, (size_t)p);
                if (is_valid(dtype, v)) {
                    if (use_kahan) {
                        double y = v - comp[g];
                        double t = sum[g] + y;
                        comp[g] = (t - sum[g]) - y;
                        sum[g] = t;
                    } else {
                        sum[g] += v;
                    }
                    ++vc[g];
                } else {
                    ++ic[g];
                }
            }
            double* o = out + (size_t)cc * G;
            for (size_t g = 0; g < G; ++g) {
                size_t denom = include_zeros ? (group_rows[g] - ic[g]) : vc[g];
                o[g] = denom > 0 ? sum[g] / (double)denom : 0.0;
            }
        }
        free(sum); free(comp); free(vc); free(ic);
    }
    free(group_rows);
    return 0;
}
