// batch_kernel.hpp -- device code of the persistent batched one-sided Jacobi (batch.hip): item checks, the
// per-item solve and the persistent kernel.  Kept apart from the host code so that the same source also builds
// for the CPU model of a workgroup under tests/simt_model.
#pragma once

#include "common.hpp"

namespace jsvd
{
    namespace
    {

        __device__ inline bool item_ok(const jsvd_batch_item &it, const jsvd_batch_buffers &b, int max_rows,
                                       int max_cols)
        {
            const int64_t m = it.rows;
            const int64_t n = it.columns;
            if (n < 1 || m < n || m > max_rows || n > max_cols || it.max_sweeps < 1 || !(it.tolerance > 0.0) ||
                (it.precision != JSVD_PRECISION_F64 && it.precision != JSVD_PRECISION_F32))
            {
                return false;
            }
            return it.a_offset >= 0 && it.a_offset + m * n <= b.a_len && it.u_offset >= 0 &&
                   it.u_offset + m * n <= b.u_len && it.sigma_offset >= 0 && it.sigma_offset + n <= b.sigma_len &&
                   it.v_offset >= 0 && it.v_offset + n * n <= b.v_len;
        }

        // Gram entries, rotation parameters (FP64) and Givens apply to A and V for P disjoint column pairs
        // (ps[t], qs[t]) of one round, by one wavefront: lane l owns rows l, l + 64, ...; each pair reduced by DPP.
        template <typename T, int P>
        __device__ inline void rotate_pairs(T *a, T *v, int lda, int ldv, int m, int n, const int (&ps)[P],
                                            const int (&qs)[P], double eps, int lane, int *sweep_rot)
        {
            T spp[P], sqq[P], spq[P];
#pragma unroll
            for (int t = 0; t < P; ++t)
            {
                spp[t] = 0;
                sqq[t] = 0;
                spq[t] = 0;
            }
            for (int i = lane; i < m; i += 64)
            {
#pragma unroll
                for (int t = 0; t < P; ++t)
                {
                    const T x = a[ps[t] * lda + i];
                    const T y = a[qs[t] * lda + i];
                    spp[t] += x * x;
                    sqq[t] += y * y;
                    spq[t] += x * y;
                }
            }
            bool go[P];
            T ct[P], st[P];
            int rotated = 0;
#pragma unroll
            for (int t = 0; t < P; ++t)
            {
                spp[t] = dev::wave_sum_dpp(spp[t]);
                sqq[t] = dev::wave_sum_dpp(sqq[t]);
                spq[t] = dev::wave_sum_dpp(spq[t]);
                double c = 1.0, sn = 0.0;
                if constexpr (sizeof(T) == 4)
                {
                    // FP32 items apply their rotations in FP32: the shorter FP64 chain (same skip test, one
                    // square root and one reciprocal square root) is exact enough there
                    go[t] = dev::rotation_params_fast(static_cast<double>(spp[t]), static_cast<double>(sqq[t]),
                                                      static_cast<double>(spq[t]), eps, c, sn);
                }
                else
                {
                    go[t] = dev::rotation_params(static_cast<double>(spp[t]), static_cast<double>(sqq[t]),
                                                 static_cast<double>(spq[t]), eps, c, sn);
                }
                ct[t] = static_cast<T>(c);
                st[t] = static_cast<T>(sn);
                rotated += go[t] ? 1 : 0;
            }
            if (rotated == 0)
            {
                return;
            }
            for (int i = lane; i < m; i += 64)
            {
#pragma unroll
                for (int t = 0; t < P; ++t)
                {
                    if (go[t])
                    {
                        T *cp = a + ps[t] * lda;
                        T *cq = a + qs[t] * lda;
                        const T x = cp[i];
                        const T y = cq[i];
                        cp[i] = ct[t] * x - st[t] * y;
                        cq[i] = st[t] * x + ct[t] * y;
                    }
                }
            }
            for (int i = lane; i < n; i += 64)
            {
#pragma unroll
                for (int t = 0; t < P; ++t)
                {
                    if (go[t])
                    {
                        T *vp = v + ps[t] * ldv;
                        T *vq = v + qs[t] * ldv;
                        const T x = vp[i];
                        const T y = vq[i];
                        vp[i] = ct[t] * x - st[t] * y;
                        vq[i] = st[t] * x + ct[t] * y;
                    }
                }
            }
            if (lane == 0)
            {
                atomicAdd(sweep_rot, rotated);
            }
        }

        template <typename T, int THREADS>
        __device__ void solve_item(const jsvd_batch_item &it, const jsvd_batch_buffers &b, T *a, T *v, int *sweep_rot,
                                   int64_t index)
        {
            constexpr int kWaves = THREADS / 64;
            const int m = it.rows;
            const int n = it.columns;
            const int lda = m + (m & 1);
            const int ldv = n + (n & 1);
            const int tid = threadIdx.x;
            const int lane = tid % 64;
            const int wave = tid / 64;
            const double eps = it.tolerance;
            // A row-major (global, FP64) -> column-major T; V = I.
            const double *src = b.a + it.a_offset;
            for (int e = tid; e < m * n; e += THREADS)
            {
                const int i = e / n;
                const int j = e % n;
                a[j * lda + i] = static_cast<T>(src[e]);
            }
            for (int e = tid; e < n * ldv; e += THREADS)
            {
                v[e] = (e / ldv == e % ldv) ? T(1) : T(0);
            }
            __syncthreads();
            const int N = n + (n & 1);
            const int rounds = (n >= 2) ? N - 1 : 0;
            int executed = 0;
            for (int s = 0; s < it.max_sweeps; ++s)
            {
                if (tid == 0)
                {
                    *sweep_rot = 0;
                }
                __syncthreads();
                for (int r = 0; r < rounds; ++r)
                {
                    // Two pairs per step (k and k + waves) with their loads, reductions and rotation parameters
                    // interleaved; every pair's own arithmetic is the same as alone.
                    for (int k = wave; k < N / 2; k += 2 * kWaves)
                    {
                        int p0 = -1, q0 = -1, p1 = -1, q1 = -1;
                        {
                            const int p = dev::circle_at_fast(k, r, N);
                            const int q = dev::circle_at_fast(N - 1 - k, r, N);
                            if (p < n && q < n)
                            {
                                p0 = p;
                                q0 = q;
                            }
                        }
                        const int k2 = k + kWaves;
                        if (k2 < N / 2)
                        {
                            const int p = dev::circle_at_fast(k2, r, N);
                            const int q = dev::circle_at_fast(N - 1 - k2, r, N);
                            if (p < n && q < n)
                            {
                                if (p0 < 0)
                                {
                                    p0 = p;
                                    q0 = q;
                                }
                                else
                                {
                                    p1 = p;
                                    q1 = q;
                                }
                            }
                        }
                        if (p1 >= 0)
                        {
                            const int ps[2] = {p0, p1};
                            const int qs[2] = {q0, q1};
                            rotate_pairs<T, 2>(a, v, lda, ldv, m, n, ps, qs, eps, lane, sweep_rot);
                        }
                        else if (p0 >= 0)
                        {
                            const int ps[1] = {p0};
                            const int qs[1] = {q0};
                            rotate_pairs<T, 1>(a, v, lda, ldv, m, n, ps, qs, eps, lane, sweep_rot);
                        }
                    }
                    __syncthreads();
                }
                executed = s + 1;
                const int rotated = *sweep_rot;
                __syncthreads(); // everyone has read the counter before the next sweep clears it
                if (rotated == 0)
                {
                    break;
                }
            }
            // sigma_j = sqrt(sum a_ij^2) in FP64 over the stored values; U = A / sigma (absolute cutoff eps).
            double *u_out = b.u + it.u_offset;
            double *s_out = b.sigma + it.sigma_offset;
            double *v_out = b.v + it.v_offset;
            for (int j = wave; j < n; j += kWaves)
            {
                const T *cj = a + j * lda;
                double acc = 0.0;
                for (int i = lane; i < m; i += 64)
                {
                    const double x = static_cast<double>(cj[i]);
                    acc += x * x;
                }
                acc = dev::wave_sum(acc);
                const double sj = sqrt(fmax(acc, 0.0));
                if (lane == 0)
                {
                    s_out[j] = sj;
                }
                for (int i = lane; i < m; i += 64)
                {
                    u_out[static_cast<int64_t>(i) * n + j] = (sj > eps) ? (static_cast<double>(cj[i]) / sj) : 0.0;
                }
            }
            for (int e = tid; e < n * n; e += THREADS)
            {
                const int i = e / n;
                const int j = e % n;
                v_out[e] = static_cast<double>(v[j * ldv + i]);
            }
            if (tid == 0)
            {
                b.sweeps[index] = executed;
            }
            __syncthreads(); // LDS / scratch reuse by the next item
        }

        template <int THREADS, int LDS_BYTES>
        __global__ __launch_bounds__(THREADS) void batch_kernel(const jsvd_batch_item *__restrict__ items,
                                                                     int64_t count, jsvd_batch_buffers b,
                                                                     unsigned long long *__restrict__ counter,
                                                                     double *__restrict__ scratch,
                                                                     int64_t scratch_per_wg, int max_rows, int max_cols)
        {
            __shared__ __align__(16) unsigned char lds[LDS_BYTES];
            __shared__ int sweep_rot;
            __shared__ long long next;
            double *my_scratch = scratch + static_cast<int64_t>(blockIdx.x) * scratch_per_wg;
            for (;;)
            {
                if (threadIdx.x == 0)
                {
                    next = static_cast<long long>(atomicAdd(counter, 1ull));
                }
                __syncthreads();
                const long long idx = next;
                __syncthreads();
                if (idx >= count)
                {
                    break; // every workgroup leaves once the queue is drained
                }
                const jsvd_batch_item it = items[idx];
                if (!item_ok(it, b, max_rows, max_cols))
                {
                    if (threadIdx.x == 0)
                    {
                        b.sweeps[idx] = -1;
                    }
                    continue;
                }
                const int m = it.rows;
                const int n = it.columns;
                const int64_t a_elems = static_cast<int64_t>(m + (m & 1)) * n;
                const int64_t v_elems = static_cast<int64_t>(n + (n & 1)) * n;
                if (it.precision == JSVD_PRECISION_F32)
                {
                    const int64_t budget = LDS_BYTES / sizeof(float);
                    float *lds_f = reinterpret_cast<float *>(lds);
                    float *scr_f = reinterpret_cast<float *>(my_scratch);
                    float *a = (a_elems <= budget) ? lds_f : scr_f;
                    const int64_t used = (a_elems <= budget) ? a_elems : 0;
                    float *v = (used + v_elems <= budget) ? lds_f + used : scr_f + a_elems;
                    solve_item<float, THREADS>(it, b, a, v, &sweep_rot, idx);
                }
                else
                {
                    const int64_t budget = LDS_BYTES / sizeof(double);
                    double *lds_d = reinterpret_cast<double *>(lds);
                    double *a = (a_elems <= budget) ? lds_d : my_scratch;
                    const int64_t used = (a_elems <= budget) ? a_elems : 0;
                    double *v = (used + v_elems <= budget) ? lds_d + used : my_scratch + a_elems;
                    solve_item<double, THREADS>(it, b, a, v, &sweep_rot, idx);
                }
            }
        }
    } // namespace
} // namespace jsvd
