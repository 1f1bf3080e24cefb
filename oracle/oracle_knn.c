/*
 * CPU restatement of the nearest-neighbour arithmetic on ennemi's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels in
 * ennemi_amd/csrc; nothing in the product links or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 *
 * It restates, by brute force, the two SciPy cKDTree calls that ennemi makes:
 *
 *   cKDTree(pts).query(pts, k=[k+1], p=np.inf)[0]
 *       ennemi/_entropy_estimators.py:39,108,142,194,240
 *       -> the (k+1)-th smallest of D(i,j) = max_d fl64(|a_d[i] - a_d[j]|)
 *          over ALL j including i itself (distance 0); +inf if n <= k.
 *
 *   cKDTree(cand).query_ball_point(q, r, p=np.inf, return_length=True)
 *       ennemi/_entropy_estimators.py:109-110,152-154,196,243-245
 *       -> #{ j : D(i,j) <= r_i }, inclusive; a negative r_i gives 0.
 *
 * Points are row-major (n, D) float64, as np.column_stack builds them
 * (ennemi/_entropy_estimators.py:93-98,134,145-146,221,233).
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp -shared -fPIC
 * (oracle/Makefile).  No FMA or reassociation may touch these loops.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static inline double chebyshev(const double* a, const double* b, int32_t D) {
    double d = 0.0;
    for (int32_t t = 0; t < D; ++t) {
        double v = fabs(a[t] - b[t]);
        if (v > d) d = v;
    }
    return d;
}

/* (k+1)-th smallest Chebyshev distance of every point to the set, self included. */
void orc_knn_eps(const double* pts, int64_t n, int32_t D, int32_t k, double* eps) {
    const int32_t K1 = k + 1;
    if (n < K1) {
        for (int64_t i = 0; i < n; ++i) eps[i] = INFINITY;
        return;
    }
#pragma omp parallel
    {
        double* list = (double*)malloc(sizeof(double) * (size_t)K1);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            for (int32_t t = 0; t < K1; ++t) list[t] = INFINITY;
            const double* p = pts + i * D;
            for (int64_t j = 0; j < n; ++j) {
                double d = chebyshev(p, pts + j * D, D);
                if (d < list[K1 - 1]) {
                    int32_t t = K1 - 1;
                    while (t > 0 && list[t - 1] > d) { list[t] = list[t - 1]; --t; }
                    list[t] = d;
                }
            }
            eps[i] = list[K1 - 1];
        }
        free(list);
    }
}

/* Inclusive max-norm ball counts of query points q against candidate set c. */
void orc_ball_count(const double* q, int64_t nq, const double* c, int64_t nc,
                    int32_t D, const double* r, int64_t* out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nq; ++i) {
        const double ri = r[i];
        int64_t cnt = 0;
        if (!(ri < 0.0)) {
            const double* p = q + i * D;
            for (int64_t j = 0; j < nc; ++j)
                if (chebyshev(p, c + j * D, D) <= ri) ++cnt;
        }
        out[i] = cnt;
    }
}
