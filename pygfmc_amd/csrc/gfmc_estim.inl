// gfmc_estim.inl -- the p_array energy estimator on the device (included at
// the end of gfmc_hip.hip, one translation unit).
//
// The reference's windowed ("forward-walking") mixed estimator
// (_energy_and_diagonal_estimates, estimators.py:268-305):
//   E_p = sum_i G_p(i) E(i + p) / sum_i G_p(i),  log G_p(i) = C[i+p] - C[i],
// C the cumulative sum of the per-step log G_n.  Streamed: a ring buffer
// holds the last p_max + 1 values of C (C[t] at slot t mod (p_max + 1)),
// and for every p a running maximum m_p with the sums num_p, den_p of
// exp(log G - m_p) E and exp(log G - m_p), rescaled when the maximum grows.
// State (doubles): [0] steps consumed, [1] C of the last step, [2] 1 once a
// non-finite log G or E was seen, [3 .. 3 + R) the ring (R = p_max + 1),
// then m[n_p], num[n_p], den[n_p].

#define FW_HDR 3

__global__ __launch_bounds__(WAVE) void k_fw_init(int n_p, int ring_len, double *st) {
    const int n = FW_HDR + ring_len + 3 * n_p;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int k = q - (FW_HDR + ring_len);
        st[q] = (k >= 0 && k < n_p) ? -INFINITY : 0.0;
    }
}

// one thread per p; rows: n_rows x (log G - log n, E) of consecutive steps
__global__ __launch_bounds__(WAVE) void k_fw_update(const int32_t *p_list, int n_p, int ring_len,
                                                    const double *rows, int n_rows,
                                                    double *st) {
    const int j = threadIdx.x;
    double *ring = st + FW_HDR;
    double *m = ring + ring_len, *num = m + n_p, *den = num + n_p;
    double t = st[0], C = st[1];
    const int p = j < n_p ? p_list[j] : 0;
    double mj = j < n_p ? m[j] : 0.0, nj = j < n_p ? num[j] : 0.0, dj = j < n_p ? den[j] : 0.0;
    bool bad = st[2] != 0.0;
    for (int r = 0; r < n_rows; ++r) {
        const double w = rows[2 * r], e = rows[2 * r + 1];
        bad = bad || !isfinite(w) || !isfinite(e);
        const double Cn = C + w;
        const int64_t ti = (int64_t)t;
        if (j < n_p && ti >= p) {
            const double lg = Cn - ring[(ti - p) % ring_len];
            if (lg > mj) {
                const double s = isinf(mj) ? 0.0 : exp(mj - lg);
                nj = nj * s + e;
                dj = dj * s + 1.0;
                mj = lg;
            } else {
                const double f = exp(lg - mj);
                nj = nj + f * e;
                dj = dj + f;
            }
        }
        __syncthreads();
        if (j == 0) ring[ti % ring_len] = Cn;
        __syncthreads();
        C = Cn;
        t = t + 1.0;
    }
    if (j < n_p) {
        m[j] = mj;
        num[j] = nj;
        den[j] = dj;
    }
    if (j == 0) {
        st[0] = t;
        st[1] = C;
        st[2] = bad ? 1.0 : 0.0;
    }
}

extern "C" {

int64_t gfmc_fw_state_doubles(int32_t n_p, int32_t p_max) {
    if (n_p < 1 || n_p > WAVE || p_max < 1) return -1;
    return FW_HDR + (int64_t)p_max + 1 + 3 * (int64_t)n_p;
}

int gfmc_fw_init(double *state, int32_t n_p, int32_t p_max, void *stream) {
    if (!state || n_p < 1 || n_p > WAVE || p_max < 1)
        return fail(GFMC_EINVAL, "gfmc_fw_init: bad argument (1 <= n_p <= 64, p_max >= 1)");
    hipLaunchKernelGGL(k_fw_init, dim3(1), dim3(WAVE), 0, (hipStream_t)stream, n_p, p_max + 1,
                       state);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_fw_update(const int32_t *p_list, int32_t n_p, int32_t p_max, const double *rows,
                   int32_t n_rows, double *state, void *stream) {
    if (!p_list || !rows || !state || n_p < 1 || n_p > WAVE || p_max < 1 || n_rows < 0)
        return fail(GFMC_EINVAL, "gfmc_fw_update: bad argument");
    if (n_rows == 0) return 0;
    hipLaunchKernelGGL(k_fw_update, dim3(1), dim3(WAVE), 0, (hipStream_t)stream, p_list, n_p,
                       p_max + 1, rows, n_rows, state);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
