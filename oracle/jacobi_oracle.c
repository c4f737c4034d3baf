/*
 * jacobi_oracle.c -- CPU restatement of the reference's cyclic one-sided (Hestenes) Jacobi SVD.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP solver's results are checked against this file: it is loaded by tests/,
 * by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and by nothing else.  The
 * product library (libjsvd.so) never links or calls it.
 *
 * Every function restates one piece of the reference (kleedaisuki/jacobi-svd-cuda) and cites the
 * file:line it follows.  Summation order is the natural row order 0..m-1 in double precision; the
 * reference's per-thread strided partial sums plus shared-memory tree (src/kernels.cu:436-464) is
 * an unknowable bit pattern on CUDA, so the oracle pins semantics, not bits.  The file is compiled
 * with -ffp-contract=off so multiply and add stay separate IEEE operations.
 *
 * What pins each function (see tests/test_oracle.py):
 *   oracle_round_robin       closed-form known answers n=4, n=5 and disjointness/coverage properties
 *   oracle_rotation          closed forms: rotated off-diagonal vanishes, c^2+s^2=1, |t|<=1,
 *                            alpha==beta -> c=s=1/sqrt(2), gamma==0 / |gamma|<=1e-300 -> skip
 *   oracle_svd               LAPACK dgesvj (SciPy-bundled OpenBLAS) agreement, closed-form 2x2 SVD,
 *                            diagonal inputs, invariants rho_A, rho_U, rho_V
 *   oracle_u_sigma           diagonal and orthonormal-column closed forms, absolute-epsilon cutoff
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/*
 * Round-robin ("circle method") schedule with element 0 fixed.
 * Follows src/kernels.cu:637-678: N = n rounded up to even; the circle starts as 0..n-1 with a
 * dummy (-1) in the last slot when n is odd; round r pairs (circle[i], circle[N-1-i]) for i < N/2,
 * dropping pairs that touch the dummy; then circle[1..N-1] rotates right by one position.
 *
 * pairs_out receives rounds*(N/2) (p,q) int pairs, counts_out[r] the number of kept pairs of
 * round r.  Returns the number of rounds (N-1), or 0 when n < 2.
 */
int oracle_round_robin(int n, int *pairs_out, int *counts_out)
{
    if (n < 2)
    {
        return 0;
    }
    const int even = (n % 2 != 0) ? n + 1 : n;
    int *circle = (int *)malloc(sizeof(int) * (size_t)even);
    for (int i = 0; i < even; ++i)
    {
        circle[i] = (i < n) ? i : -1;
    }
    const int half = even / 2;
    for (int r = 0; r < even - 1; ++r)
    {
        int kept = 0;
        for (int i = 0; i < half; ++i)
        {
            const int lhs = circle[i];
            const int rhs = circle[even - 1 - i];
            if (lhs >= 0 && rhs >= 0)
            {
                pairs_out[2 * ((size_t)r * half + kept) + 0] = lhs;
                pairs_out[2 * ((size_t)r * half + kept) + 1] = rhs;
                ++kept;
            }
        }
        counts_out[r] = kept;
        const int last = circle[even - 1];
        for (int i = even - 1; i > 1; --i)
        {
            circle[i] = circle[i - 1];
        }
        circle[1] = last;
    }
    free(circle);
    return even - 1;
}

/*
 * Rotation parameters and skip test, src/kernels.cu:485-521.
 *   rhs = eps * sqrt(max(app,0) * max(aqq,0)); rotate iff |apq| > rhs and |apq| > 1e-300;
 *   tau = (aqq - app) / (2 apq); t = 1/(tau + sqrt(1+tau^2)) for tau >= 0, else -1/(-tau + sqrt(1+tau^2));
 *   c = 1/sqrt(1+t^2); s = t c.  Without rotation c = 1, s = 0.
 * Returns 1 when the pair rotates.
 */
int oracle_rotation(double app, double aqq, double apq, double eps, double *c, double *s)
{
    const double rhs = eps * sqrt(fmax(app, 0.0) * fmax(aqq, 0.0));
    if (fabs(apq) > rhs && fabs(apq) > 1.0e-300)
    {
        const double tau = (aqq - app) / (2.0 * apq);
        const double root = sqrt(1.0 + tau * tau);
        const double t = (tau >= 0.0) ? (1.0 / (tau + root)) : (-1.0 / (-tau + root));
        const double cosine = 1.0 / sqrt(1.0 + t * t);
        *c = cosine;
        *s = t * cosine;
        return 1;
    }
    *c = 1.0;
    *s = 0.0;
    return 0;
}

/*
 * Pair statistics of one column pair, src/kernels.cu:412-472 (semantics; row order 0..m-1).
 * acol is column-major with leading dimension lda.
 */
void oracle_pair_stats(const double *acol, int64_t lda, int64_t m, int p, int q,
                       double *app, double *aqq, double *apq)
{
    const double *cp = acol + (size_t)p * (size_t)lda;
    const double *cq = acol + (size_t)q * (size_t)lda;
    double spp = 0.0, sqq = 0.0, spq = 0.0;
    for (int64_t i = 0; i < m; ++i)
    {
        const double x = cp[i];
        const double y = cq[i];
        spp += x * x;
        sqq += y * y;
        spq += x * y;
    }
    *app = spp;
    *aqq = sqq;
    *apq = spq;
}

/*
 * Givens apply to one column pair of a column-major matrix, src/kernels.cu:534-575:
 *   x' = c x - s y,  y' = s x + c y.
 */
void oracle_apply(double *col, int64_t ld, int64_t rows, int p, int q, double c, double s)
{
    double *cp = col + (size_t)p * (size_t)ld;
    double *cq = col + (size_t)q * (size_t)ld;
    for (int64_t i = 0; i < rows; ++i)
    {
        const double x = cp[i];
        const double y = cq[i];
        cp[i] = c * x - s * y;
        cq[i] = s * x + c * y;
    }
}

/*
 * sigma and U extraction, src/kernels.cu:586-629: sigma_j = sqrt(max(sum_i a_ij^2, 0));
 * u_ij = a_ij / sigma_j when sigma_j > eps (absolute cutoff with the solver's epsilon), else 0.
 * acol is column-major (lda); u_rm is written row-major m x n.
 */
void oracle_u_sigma(const double *acol, int64_t lda, int64_t m, int64_t n, double eps,
                    double *u_rm, double *sigma)
{
    for (int64_t j = 0; j < n; ++j)
    {
        const double *cj = acol + (size_t)j * (size_t)lda;
        double acc = 0.0;
        for (int64_t i = 0; i < m; ++i)
        {
            acc += cj[i] * cj[i];
        }
        const double sj = sqrt(fmax(acc, 0.0));
        sigma[j] = sj;
        for (int64_t i = 0; i < m; ++i)
        {
            u_rm[(size_t)i * (size_t)n + (size_t)j] = (sj > eps) ? (cj[i] / sj) : 0.0;
        }
    }
}

/*
 * Full solve, src/kernels.cu:844-1004 (driver) with 887-981 (sweep/round loop):
 *   V = I; for sweep < max_sweeps: for each round: stats, rotation, apply to A and V;
 *   sweeps = sweep + 1; stop after the first sweep in which no pair rotated.
 * Non-convergence at max_sweeps is not an error (sweeps == max_sweeps).
 * Input row-major m x n (m >= n); outputs row-major U (m x n), sigma (n), V (n x n).
 * rot_per_sweep (optional, max_sweeps ints) receives the number of rotated pairs per sweep,
 * offmax_per_sweep (optional) the largest |apq|/sqrt(app aqq) seen in each sweep.
 * nthreads > 1 parallelises the disjoint pairs of each round with OpenMP (barrier per round).
 * Returns 0 on success, 1 on invalid arguments (src/kernels.cu:807-833).
 */
int oracle_svd_ex(const double *a_rm, int64_t m, int64_t n, double eps, int max_sweeps, int nthreads,
                  double *u_rm, double *sigma, double *v_rm, int *sweeps_out,
                  int *rot_per_sweep, double *offmax_per_sweep)
{
    if (m <= 0 || n <= 0 || m < n || !(eps > 0.0) || max_sweeps <= 0)
    {
        return 1;
    }
    const int64_t lda = m;
    double *acol = (double *)malloc(sizeof(double) * (size_t)(m * n));
    double *vcol = (double *)malloc(sizeof(double) * (size_t)(n * n));
    for (int64_t i = 0; i < m; ++i)
    {
        for (int64_t j = 0; j < n; ++j)
        {
            acol[(size_t)j * (size_t)lda + (size_t)i] = a_rm[(size_t)i * (size_t)n + (size_t)j];
        }
    }
    for (int64_t j = 0; j < n; ++j)
    {
        for (int64_t i = 0; i < n; ++i)
        {
            vcol[(size_t)j * (size_t)n + (size_t)i] = (i == j) ? 1.0 : 0.0;
        }
    }

    const int even = (n % 2 != 0) ? (int)n + 1 : (int)n;
    const int half = even / 2;
    int rounds = 0;
    int *pairs = NULL;
    int *counts = NULL;
    if (n >= 2)
    {
        pairs = (int *)malloc(sizeof(int) * 2 * (size_t)(even - 1) * (size_t)half);
        counts = (int *)malloc(sizeof(int) * (size_t)(even - 1));
        rounds = oracle_round_robin((int)n, pairs, counts);
    }

#ifdef _OPENMP
    if (nthreads > 0)
    {
        omp_set_num_threads(nthreads);
    }
#else
    (void)nthreads;
#endif

    int executed = 0;
    for (int sweep = 0; sweep < max_sweeps; ++sweep)
    {
        int rotated = 0;
        double offmax = 0.0;
        for (int r = 0; r < rounds; ++r)
        {
            const int count = counts[r];
            const int *rp = pairs + 2 * (size_t)r * (size_t)half;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : rotated) reduction(max : offmax) if (nthreads > 1)
#endif
            for (int k = 0; k < count; ++k)
            {
                const int p = rp[2 * k];
                const int q = rp[2 * k + 1];
                double app, aqq, apq, c, s;
                oracle_pair_stats(acol, lda, m, p, q, &app, &aqq, &apq);
                const double denom = sqrt(fmax(app, 0.0) * fmax(aqq, 0.0));
                if (denom > 0.0)
                {
                    const double cosine = fabs(apq) / denom;
                    if (cosine > offmax)
                    {
                        offmax = cosine;
                    }
                }
                if (oracle_rotation(app, aqq, apq, eps, &c, &s))
                {
                    rotated += 1;
                    oracle_apply(acol, lda, m, p, q, c, s);
                    oracle_apply(vcol, n, n, p, q, c, s);
                }
            }
        }
        if (rot_per_sweep)
        {
            rot_per_sweep[sweep] = rotated;
        }
        if (offmax_per_sweep)
        {
            offmax_per_sweep[sweep] = offmax;
        }
        executed = sweep + 1;
        if (rotated == 0)
        {
            break;
        }
    }

    oracle_u_sigma(acol, lda, m, n, eps, u_rm, sigma);
    for (int64_t i = 0; i < n; ++i)
    {
        for (int64_t j = 0; j < n; ++j)
        {
            v_rm[(size_t)i * (size_t)n + (size_t)j] = vcol[(size_t)j * (size_t)n + (size_t)i];
        }
    }
    *sweeps_out = executed;
    free(acol);
    free(vcol);
    free(pairs);
    free(counts);
    return 0;
}

int oracle_svd(const double *a_rm, int64_t m, int64_t n, double eps, int max_sweeps, int nthreads,
               double *u_rm, double *sigma, double *v_rm, int *sweeps_out)
{
    return oracle_svd_ex(a_rm, m, n, eps, max_sweeps, nthreads, u_rm, sigma, v_rm, sweeps_out, NULL, NULL);
}

/* Number of OpenMP threads the oracle would use (1 without OpenMP). */
int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
