// gfmc_train.inl -- training kernels of the RBM guide (included at the end of
// gfmc_hip.hip, one translation unit).
//
// Loss: the variance of the local energy over n samples x_i that GFMC
// generated,
//   L = (1/n) sum_i (E_i - Ebar)^2,   E_i = sum_k rates_k(x_i),
// with rates_k = H_k psi(x^k) / psi(x) the guide-transformed rates the
// propagation kernels use.  Its gradient (d_i = E_i - Ebar, optionally
// clamped to +-clip*sqrt(L); O = d log psi / dp):
//   dL/dp = sum_i g_i sum_{k>=1} rates_k(x_i) (O_p(x^k_i) - O_p(x_i)),
//   g_i = 2 d_i / n.
// For log psi = a.x + sum_h lncosh(theta_h), theta = b + W x, with
// c_ik = g_i rates_k(x_i), C_i = sum_k c_ik, t = tanh(theta(x_i)) and
// t'_k = tanh(theta(x^k_i)):
//   dL/db_h   = sum_i B_ih,    B_ih = sum_k c_ik t'_kh - C_i t_h
//   dL/dW_hm  = sum_i x_im B_ih  -  2 sum_i x_im sum_{k: m in S_k} c_ik t'_kh
//   dL/da_m   = -2 sum_i x_im sum_{k: m in S_k} c_ik
// (S_k: the sites slot k flips).  The kernels:
//   k_rbm_theta_mfma  theta = X W^T + b          BF16 MFMA, FP32 accumulation
//   k_vmin_coef       mean, variance, g_i         one workgroup, fixed order
//   k_vmin_moves      B, and the per-site sums    VALU; lanes = hidden units
//   k_rbm_bx_mfma     sum_i B_ih x_im             BF16 MFMA on a hi/lo split of B
//   k_vmin_reduce     partials -> flat gradient   fixed order
//   k_adam            Adam on the flat FP32 parameters
//   k_rbm_pack/_cs    FP32 parameters -> the problem's BF16 rows and column sums
// Every reduction runs in a fixed order: the gradient is bit-reproducible.

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define BF16_ONE ((short)0x3F80)
#define BF16_MINUS_ONE ((short)0xBF80)
#define VM_SPW 16     // samples per wavefront in k_vmin_moves
#define BX_KS 256     // samples per split in k_rbm_bx_mfma

__device__ __forceinline__ uint32_t bf16_rne_bits(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ float tanh_acc(float x) {
    const float em = expm1f(-2.0f * fabsf(x));
    return copysignf(-em / (2.0f + em), x);
}

__device__ __forceinline__ float lane_bcast_f(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// theta[i][h] = b_h + sum_k W_hk x_ik.  One wavefront per 32 x 32 tile
// (samples x hidden units) of v_mfma_f32_32x32x8_bf16: lane l holds
// A[i = l%32][k = 4(l/32)+q] = x_ik and B[k = 4(l/32)+q][h = l%32] = W_hk,
// and D[i][h] in register r at i = 8(r/4) + 4(l/32) + r%4, h = l%32.
__global__ __launch_bounds__(256) void k_rbm_theta_mfma(DevProblem P, int64_t n,
                                                        const uint32_t *spins, float *theta) {
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const int ht = (P.H + 31) / 32;
    const int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    if (tile >= (n + 31) / 32 * ht) return;
    const int64_t i0 = tile / ht * 32;
    const int h0 = (int)(tile % ht) * 32;
    const int col = lane & 31, half = lane >> 5;
    const int64_t ia = i0 + col;
    const int hb = h0 + col;
    const int Np = P.w_words * 2 / P.H;
    const uint32_t *sp = spins + (ia < n ? ia : 0) * P.words;
    f32x16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int k0 = 0; k0 < P.N; k0 += 8) {
        const int k = k0 + 4 * half;
        const uint32_t bits = (ia < n && k < P.N) ? sp[k >> 5] >> (k & 31) : 0u;
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool kv = k + q < P.N;
            a[q] = (ia < n && kv) ? (((bits >> q) & 1u) ? BF16_ONE : BF16_MINUS_ONE) : (short)0;
            b[q] = (hb < P.H && kv)
                       ? (short)(__float_as_uint(bf16_at(P.rbm_W, hb * Np + k + q)) >> 16)
                       : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    if (hb >= P.H) return;
    const float bias = P.rbm_b[hb];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t i = i0 + 8 * (r / 4) + 4 * half + (r % 4);
        if (i < n) theta[i * P.H + hb] = acc[r] + bias;
    }
}

// mean and variance of e (blocked per thread, tree over the workgroup) and
// g_i = 2 clamp(e_i - mean) / n; out2 = (mean, variance)
__global__ __launch_bounds__(RECONF_THREADS) void k_vmin_coef(int64_t n, const double *e,
                                                              float clip, float *g,
                                                              double *out2) {
    __shared__ double red[RECONF_THREADS];
    const int t = threadIdx.x;
    const int64_t per = (n + RECONF_THREADS - 1) / RECONF_THREADS;
    const int64_t lo = t * per, hi = lo + per < n ? lo + per : n;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s += e[i];
    red[t] = s;
    __syncthreads();
    for (int w = RECONF_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) red[t] = red[t] + red[t + w];
        __syncthreads();
    }
    const double mean = red[0] / (double)n;
    __syncthreads();
    s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s += (e[i] - mean) * (e[i] - mean);
    red[t] = s;
    __syncthreads();
    for (int w = RECONF_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) red[t] = red[t] + red[t + w];
        __syncthreads();
    }
    const double var = red[0] / (double)n;
    const double lim = clip > 0.0f ? (double)clip * sqrt(var) : INFINITY;
    for (int64_t i = t; i < n; i += RECONF_THREADS) {
        double d = e[i] - mean;
        d = d > lim ? lim : (d < -lim ? -lim : d);
        g[i] = (float)(2.0 * d / (double)n);
    }
    if (t == 0) {
        out2[0] = mean;
        out2[1] = var;
    }
}

// LDS bytes of k_vmin_moves with vw wavefronts: the W^T slice [N][64] FP32,
// then per wavefront acc [N][64] FP32, A [N] FP32 and x [N] int8
__host__ __device__ inline size_t vmin_wave_bytes(int N) {
    return (size_t)N * 64 * 4 + (size_t)N * 4 + (size_t)(N + 15) / 16 * 16;
}
__host__ __device__ inline size_t vmin_lds_bytes(int N, int vw) {
    return (size_t)N * 64 * 4 + vw * vmin_wave_bytes(N);
}

// grid (ceil(H / 64), ceil(n_cw / vw)), vw wavefronts per workgroup; wave w of
// workgroup y owns samples [cw * 64, cw * 64 + 64), cw = y * vw + w, and lanes
// own the hidden units h0 + lane of tile x.  Per sample, the slots with a
// nonzero coefficient are visited in slot order (ballot + readlane).
__global__ __launch_bounds__(256) void k_vmin_moves(DevProblem P, int64_t n,
                                                    const uint32_t *spins, const float *theta,
                                                    const double *rates, const float *g,
                                                    float *Bout, float *Dpart, float *bpart,
                                                    float *apart) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const int vw = blockDim.x / WAVE;
    const int N = P.N, H = P.H, Np = P.w_words * 2 / P.H;
    const int h0 = blockIdx.x * WAVE, h = h0 + lane;
    const bool hv = h < H;
    float *Wt = (float *)lds;
    unsigned char *wb = lds + (size_t)N * 64 * 4 + (size_t)wave * vmin_wave_bytes(N);
    float *acc = (float *)wb;
    float *A = acc + (size_t)N * 64;
    signed char *x = (signed char *)(A + N);
    for (int t = threadIdx.x; t < N * 64; t += blockDim.x) {
        const int m = t / 64, c = t % 64;
        Wt[t] = h0 + c < H ? bf16_at(P.rbm_W, (h0 + c) * Np + m) : 0.0f;
    }
    for (int t = lane; t < N * 64; t += WAVE) acc[t] = 0.0f;
    for (int t = lane; t < N; t += WAVE) A[t] = 0.0f;
    __syncthreads();
    const int64_t n_cw = (n + VM_SPW - 1) / VM_SPW;
    const int64_t cw = (int64_t)blockIdx.y * vw + wave;
    if (cw >= n_cw) return;
    const int64_t i_end = (cw + 1) * VM_SPW < n ? (cw + 1) * VM_SPW : n;
    const int64_t M = P.n_off + 1;
    float gb = 0.0f;
    for (int64_t i = cw * VM_SPW; i < i_end; ++i) {
        const uint32_t *sp = spins + i * P.words;
        for (int k = lane; k < N; k += WAVE) x[k] = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1 : -1;
        wave_sync();
        const float th = hv ? theta[i * H + h] : 0.0f;
        const float gi = g[i];
        float B = 0.0f, C = 0.0f;
        if (gi != 0.0f) {
            for (int c = 0; c < P.n_off; c += WAVE) {
                const int s = c + lane;
                const float cl = s < P.n_off ? gi * (float)rates[i * M + 1 + s] : 0.0f;
                unsigned long long mask = __ballot(cl != 0.0f);
                while (mask) {
                    const int bl = __builtin_ctzll(mask);
                    mask &= mask - 1ull;
                    const float cs = lane_bcast_f(cl, bl);
                    int si, sj;
                    slot_sites(P, c + bl, &si, &sj);
                    const float xi = (float)x[si];
                    float d = -2.0f * xi * Wt[si * 64 + lane];
                    float xj = 0.0f;
                    if (sj >= 0) {
                        xj = (float)x[sj];
                        d += -2.0f * xj * Wt[sj * 64 + lane];
                    }
                    const float ct = cs * tanh_acc(th + d);
                    B += ct;
                    C += cs;
                    acc[si * 64 + lane] += -2.0f * xi * ct;
                    if (sj >= 0) acc[sj * 64 + lane] += -2.0f * xj * ct;
                    if (lane == 0) {
                        A[si] += -2.0f * xi * cs;
                        if (sj >= 0) A[sj] += -2.0f * xj * cs;
                    }
                }
            }
        }
        B -= C * tanh_acc(th);
        if (hv) Bout[i * H + h] = B;
        gb += B;
        wave_sync();
    }
    wave_sync();
    if (hv) {
        float *Dp = Dpart + cw * (int64_t)N * H;
        for (int m = 0; m < N; ++m) Dp[(int64_t)m * H + h] = acc[m * 64 + lane];
        bpart[cw * H + h] = gb;
    }
    if (blockIdx.x == 0)
        for (int m = lane; m < N; m += WAVE) apart[cw * N + m] = A[m];
}

// G1part[split][h][m] = sum_{i in split} B_ih x_im.  One wavefront per
// (32 x 32 tile of (h, m), split of BX_KS samples); A = B^T as hi + lo BF16
// halves (B = hi + lo + O(2^-17 |B|)), B operand = x (exact in BF16).
__global__ __launch_bounds__(256) void k_rbm_bx_mfma(DevProblem P, int64_t n,
                                                     const uint32_t *spins, const float *Bm,
                                                     float *G1part) {
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const int nh = (P.H + 31) / 32, nm = (P.N + 31) / 32;
    const int64_t n_split = (n + BX_KS - 1) / BX_KS;
    const int64_t t = (int64_t)blockIdx.x * 4 + wave;
    if (t >= n_split * nh * nm) return;
    const int64_t split = t / (nh * nm);
    const int r0 = (int)(t % (nh * nm));
    const int h0 = r0 / nm * 32, m0 = r0 % nm * 32;
    const int col = lane & 31, half = lane >> 5;
    const int ha = h0 + col, mb = m0 + col;
    const int64_t i_begin = split * BX_KS;
    const int64_t i_end = i_begin + BX_KS < n ? i_begin + BX_KS : n;
    f32x16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int64_t i0 = i_begin; i0 < i_end; i0 += 8) {
        bf16x4 ahi, alo, bx;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t i = i0 + 4 * half + q;
            const bool iv = i < i_end;
            const float v = (iv && ha < P.H) ? Bm[i * P.H + ha] : 0.0f;
            const uint32_t hb = bf16_rne_bits(v);
            const float lo = v - __uint_as_float(hb << 16);
            ahi[q] = (short)hb;
            alo[q] = (short)bf16_rne_bits(lo);
            bx[q] = (iv && mb < P.N)
                        ? (((spins[i * P.words + (mb >> 5)] >> (mb & 31)) & 1u) ? BF16_ONE
                                                                                 : BF16_MINUS_ONE)
                        : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ahi, bx, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(alo, bx, acc, 0, 0, 0);
    }
    if (mb >= P.N) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int h = h0 + 8 * (r / 4) + 4 * half + (r % 4);
        if (h < P.H) G1part[(split * P.H + h) * P.N + mb] = acc[r];
    }
}

// flat gradient [W (H x N), a (N), b (H)] from the partials, in fixed order.
// Threads of the W part run m-major (consecutive threads: consecutive h).
__global__ __launch_bounds__(256) void k_vmin_reduce(int N, int H, int64_t n_cw, int64_t n_split,
                                                     const float *G1part, const float *Dpart,
                                                     const float *bpart, const float *apart,
                                                     float *grad) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nW = (int64_t)H * N;
    if (q < nW) {
        const int m = (int)(q / H), h = (int)(q % H);
        float s = 0.0f;
        for (int64_t sp = 0; sp < n_split; ++sp) s += G1part[(sp * H + h) * N + m];
        for (int64_t c = 0; c < n_cw; ++c) s += Dpart[(c * N + m) * H + h];
        grad[(int64_t)h * N + m] = s;
    } else if (q < nW + N) {
        const int m = (int)(q - nW);
        float s = 0.0f;
        for (int64_t c = 0; c < n_cw; ++c) s += apart[c * N + m];
        grad[q] = s;
    } else if (q < nW + N + H) {
        const int h = (int)(q - nW - N);
        float s = 0.0f;
        for (int64_t c = 0; c < n_cw; ++c) s += bpart[c * H + h];
        grad[q] = s;
    }
}

// Adam (Kingma & Ba, ICLR 2015) in the form of torch.optim.Adam:
// m <- lerp(m, g, 1 - b1), v <- b2 v + (1 - b2) g^2,
// p <- p - (lr / bc1) m / (sqrt(v) / sqrt(bc2) + eps)
__global__ __launch_bounds__(256) void k_adam(int64_t n, float *p, const float *g, float *m,
                                              float *v, float step_size, float bc2_sqrt, float b1,
                                              float b2, float eps) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = m[q] + (1.0f - b1) * (gq - m[q]);
    const float vq = b2 * v[q] + (1.0f - b2) * (gq * gq);
    m[q] = mq;
    v[q] = vq;
    p[q] = p[q] - step_size * (mq / (sqrtf(vq) / bc2_sqrt + eps));
}

// FP32 flat parameters -> the problem's BF16 weight rows (round to nearest
// even), a, b; thread per packed word, then per a/b entry
__global__ __launch_bounds__(256) void k_rbm_pack(int N, int H, int Np, const float *flat,
                                                  uint32_t *W, float *a, float *b) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t words = (int64_t)H * Np / 2;
    if (q < words) {
        const int h = (int)(q / (Np / 2)), k = (int)(q % (Np / 2)) * 2;
        const uint32_t lo = k < N ? bf16_rne_bits(flat[(int64_t)h * N + k]) : 0u;
        const uint32_t hi = k + 1 < N ? bf16_rne_bits(flat[(int64_t)h * N + k + 1]) : 0u;
        W[q] = lo | (hi << 16);
    } else if (q < words + N) {
        const int k = (int)(q - words);
        a[k] = flat[(int64_t)H * N + k];
    } else if (q < words + N + H) {
        const int h = (int)(q - words - N);
        b[h] = flat[(int64_t)H * N + N + h];
    }
}

// column sums sum_h W_hk of the BF16 weights, h ascending in FP64 (the order
// gfmc_problem_create uses)
__global__ __launch_bounds__(256) void k_rbm_colsum(int N, int H, int Np, const uint32_t *W,
                                                    float *cs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    double s = 0.0;
    for (int h = 0; h < H; ++h) s += (double)bf16_at(W, h * Np + k);
    cs[k] = (float)s;
}

// mean over every orbit of tied parameters, in place: orbit o holds the flat
// indices members[off[o] .. off[o + 1]), added in that order in FP64; indices
// outside [0, n) are skipped
__global__ __launch_bounds__(256) void k_orbit_mean(int64_t n, int32_t n_orb, const int32_t *off,
                                                    const int32_t *members, float *v) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_orb) return;
    const int a = off[o], b = off[o + 1];
    if (b <= a) return;
    double s = 0.0;
    int cnt = 0;
    for (int t = a; t < b; ++t) {
        const int q = members[t];
        if (q < 0 || q >= n) continue;
        s += (double)v[q];
        ++cnt;
    }
    if (cnt == 0) return;
    const float m = (float)(s / (double)cnt);
    for (int t = a; t < b; ++t) {
        const int q = members[t];
        if (q >= 0 && q < n) v[q] = m;
    }
}

// v[q] = sum of v_full over orbit q's members (in members order, FP64)
__global__ __launch_bounds__(256) void k_orbit_sum(int64_t n, int32_t n_orb, const int32_t *off,
                                                   const int32_t *members, const float *v,
                                                   float *out) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_orb) return;
    double s = 0.0;
    for (int t = off[o]; t < off[o + 1]; ++t) {
        const int q = members[t];
        if (q >= 0 && q < n) s += (double)v[q];
    }
    out[o] = (float)s;
}

// v_full[p] += alpha * x[orbit of p]
__global__ __launch_bounds__(256) void k_orbit_axpy(int64_t n, int32_t n_orb, const int32_t *off,
                                                    const int32_t *members, const float *x,
                                                    float alpha, float *v) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_orb) return;
    const float d = alpha * x[o];
    for (int t = off[o]; t < off[o + 1]; ++t) {
        const int q = members[t];
        if (q >= 0 && q < n) v[q] += d;
    }
}

// Orbit-summed log-derivatives of the RBM, O[i][o] = sum over the members p
// of orbit o of d log psi(x_i) / d p: tanh(theta_ih) x_ik for W_hk, x_ik for
// a_k, tanh(theta_ih) for b_h; thread per (orbit, sample)
__global__ __launch_bounds__(256) void k_rbm_orbit_O(DevProblem P, int64_t n, const uint32_t *spins,
                                                     const float *theta, int32_t n_orb,
                                                     const int32_t *off, const int32_t *members,
                                                     float *O) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (o >= n_orb || i >= n) return;
    const int N = P.N, H = P.H;
    const int64_t HN = (int64_t)H * N;
    const uint32_t *sp = spins + i * P.words;
    const float *th = theta + i * H;
    double s = 0.0;
    for (int t = off[o]; t < off[o + 1]; ++t) {
        const int64_t q = members[t];
        if (q < HN) {
            const int h = (int)(q / N), k = (int)(q % N);
            const float xk = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1.0f : -1.0f;
            s += (double)(tanhf(th[h]) * xk);
        } else if (q < HN + N) {
            const int k = (int)(q - HN);
            s += ((sp[k >> 5] >> (k & 31)) & 1u) ? 1.0 : -1.0;
        } else if (q < HN + N + H) {
            s += (double)tanhf(th[q - HN - N]);
        }
    }
    O[i * n_orb + o] = (float)s;
}

// natural-gradient solve (stochastic reconfiguration): column means of O
// (n x m), then the covariance S = (1/n) sum_i (O_i - mean)(O_i - mean)^T in
// 16 x 16 tiles (FP64 accumulation, samples in order), then conjugate
// gradients on (S + eps_rel diag S + eps_abs) d = g in one workgroup
#define SR_T 16
#define SR_MC 32   // sample chunks of the column means
__global__ __launch_bounds__(256) void k_sr_mean_part(int64_t n, int m, const float *O,
                                                      double *part) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (q >= m) return;
    const int64_t per = (n + SR_MC - 1) / SR_MC, i0 = c * per, i1 = i0 + per < n ? i0 + per : n;
    double s = 0.0;
    for (int64_t i = i0; i < i1; ++i) s += (double)O[i * m + q];
    part[(int64_t)c * m + q] = s;
}

__global__ __launch_bounds__(256) void k_sr_mean(int64_t n, int m, const double *part,
                                                 double *mean) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    double s = 0.0;
    for (int c = 0; c < SR_MC; ++c) s += part[(int64_t)c * m + q];
    mean[q] = s / (double)n;
}

__global__ __launch_bounds__(SR_T * SR_T) void k_sr_cov(int64_t n, int m, const float *O,
                                                       const double *mean, double *S) {
    __shared__ double a[SR_T][SR_T], b[SR_T][SR_T];
    const int tx = threadIdx.x % SR_T, ty = threadIdx.x / SR_T;
    const int q1 = blockIdx.y * SR_T + ty, q2 = blockIdx.x * SR_T + tx;
    double acc = 0.0;
    for (int64_t i0 = 0; i0 < n; i0 += SR_T) {
        // a[r][c]: sample i0 + r, column blockIdx.y*T + c; b likewise for x
        const int64_t ia = i0 + ty;
        const int ca = blockIdx.y * SR_T + tx, cb = blockIdx.x * SR_T + tx;
        a[ty][tx] = (ia < n && ca < m) ? (double)O[ia * m + ca] - mean[ca] : 0.0;
        b[ty][tx] = (ia < n && cb < m) ? (double)O[ia * m + cb] - mean[cb] : 0.0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < SR_T; ++r) acc += a[r][ty] * b[r][tx];
        __syncthreads();
    }
    if (q1 < m && q2 < m) S[(int64_t)q1 * m + q2] = acc / (double)n;
}

__device__ double sr_block_sum(double v, double *sh) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x / WAVE); ++q) t += sh[q];
    __syncthreads();
    return t;
}

// y = (S + eps_rel diag S + eps_abs) x with S symmetric: y_r = sum_c S[c][r]
// x_c, so thread r reads row c of S at column r (coalesced, every load
// independent of the running sum's); the threads in two halves over the
// columns [0, ceil(m / 2)) and the rest, each summing its columns in
// ascending order, the halves' sums (part, 2 m doubles) added in that order
__device__ void sr_apply(int m, const double *S, double eps_rel, double eps_abs, const double *x,
                         double *y, double *part) {
    const int hb = blockDim.x >> 1, half = threadIdx.x >= hb ? 1 : 0, lr = threadIdx.x - half * hb;
    const int cm = (m + 1) >> 1, c0 = half ? cm : 0, c1 = half ? m : cm;
    for (int r = lr; r < m; r += hb) {
        double acc = 0.0;
        // unrolled so that eight rows' loads are in flight at once; the
        // sum keeps its ascending order
#pragma unroll 8
        for (int c = c0; c < c1; ++c) acc += S[(int64_t)c * m + r] * x[c];
        part[(int64_t)half * m + r] = acc;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < m; r += blockDim.x)
        y[r] = (part[r] + part[m + r]) + (eps_rel * S[(int64_t)r * m + r] + eps_abs) * x[r];
    __syncthreads();
}

// warm: start from the x the workspace holds (the previous call's unscaled
// solution) with r = g - A x, else from x = 0
__global__ __launch_bounds__(1024) void k_sr_cg(int m, const double *S, double eps_rel,
                                               double eps_abs, const float *g, int max_iter,
                                               double tol, double max_norm, int warm, double *x,
                                               double *r, double *pv, double *Ap, double *part,
                                               float *delta) {
    __shared__ double sh[16];
    if (warm) {
        sr_apply(m, S, eps_rel, eps_abs, x, Ap, part);
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            r[q] = (double)g[q] - Ap[q];
            pv[q] = r[q];
        }
    } else {
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            x[q] = 0.0;
            r[q] = (double)g[q];
            pv[q] = r[q];
        }
    }
    __syncthreads();
    double rr = 0.0;
    for (int q = threadIdx.x; q < m; q += blockDim.x) rr += r[q] * r[q];
    rr = sr_block_sum(rr, sh);
    const double stop = tol * tol * rr;
    for (int it = 0; it < max_iter && rr > stop && rr > 0.0; ++it) {
        sr_apply(m, S, eps_rel, eps_abs, pv, Ap, part);
        double pAp = 0.0;
        for (int q = threadIdx.x; q < m; q += blockDim.x) pAp += pv[q] * Ap[q];
        pAp = sr_block_sum(pAp, sh);
        if (!(pAp > 0.0)) break;
        const double alpha = rr / pAp;
        double rr2 = 0.0;
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            x[q] += alpha * pv[q];
            r[q] -= alpha * Ap[q];
            rr2 += r[q] * r[q];
        }
        rr2 = sr_block_sum(rr2, sh);
        const double beta = rr2 / rr;
        rr = rr2;
        for (int q = threadIdx.x; q < m; q += blockDim.x) pv[q] = r[q] + beta * pv[q];
        __syncthreads();
    }
    // trust region: the step's rms change of log psi over the samples,
    // sqrt(x^T S x), at most max_norm (0: no limit)
    double scale = 1.0;
    if (max_norm > 0.0) {
        sr_apply(m, S, 0.0, 0.0, x, Ap, part);
        double xSx = 0.0;
        for (int q = threadIdx.x; q < m; q += blockDim.x) xSx += x[q] * Ap[q];
        xSx = sr_block_sum(xSx, sh);
        if (xSx > max_norm * max_norm) scale = max_norm / sqrt(xSx);
    }
    for (int q = threadIdx.x; q < m; q += blockDim.x) delta[q] = (float)(x[q] * scale);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct VminLayout {
    int64_t n_cw, n_split;
    size_t rates, e_loc, g, theta, B, Dpart, bpart, apart, G1part, total;
};

static size_t align256(size_t v) { return (v + 255) / 256 * 256; }

static VminLayout vmin_layout(const gfmc_problem *p, int64_t n) {
    VminLayout L;
    const int64_t N = p->dp.N, H = p->dp.H, M = p->dp.n_off + 1;
    L.n_cw = (n + VM_SPW - 1) / VM_SPW;
    L.n_split = (n + BX_KS - 1) / BX_KS;
    size_t o = 0;
    L.rates = o;
    o = align256(o + (size_t)n * M * 8);
    L.e_loc = o;
    o = align256(o + (size_t)n * 8);
    L.g = o;
    o = align256(o + (size_t)n * 4);
    L.theta = o;
    o = align256(o + (size_t)n * H * 4);
    L.B = o;
    o = align256(o + (size_t)n * H * 4);
    L.Dpart = o;
    o = align256(o + (size_t)L.n_cw * N * H * 4);
    L.bpart = o;
    o = align256(o + (size_t)L.n_cw * H * 4);
    L.apart = o;
    o = align256(o + (size_t)L.n_cw * N * 4);
    L.G1part = o;
    o = align256(o + (size_t)L.n_split * H * N * 4);
    L.total = o;
    return L;
}

static int vmin_waves(int N) {
    for (int vw = 4; vw >= 1; --vw)
        if (vmin_lds_bytes(N, vw) <= 160 * 1024) return vw;
    return 0;
}

static int check_rbm(const char *fn, const gfmc_problem *p) {
    if (!p) return fail(GFMC_EINVAL, std::string(fn) + ": NULL problem");
    if (p->dp.guide != GFMC_GUIDE_RBM)
        return fail(GFMC_EINVAL, std::string(fn) + ": the problem's guide is not an RBM");
    return 0;
}

extern "C" {

int64_t gfmc_rbm_n_params(const gfmc_problem *p) {
    if (check_rbm("gfmc_rbm_n_params", p)) return -1;
    return (int64_t)p->dp.H * p->dp.N + p->dp.N + p->dp.H;
}

int gfmc_rbm_set_params(gfmc_problem *p, const float *flat, void *stream) {
    int rc = check_rbm("gfmc_rbm_set_params", p);
    if (rc) return rc;
    if (!flat) return fail(GFMC_EINVAL, "gfmc_rbm_set_params: NULL parameters");
    const int N = p->dp.N, H = p->dp.H, Np = p->dp.w_words * 2 / H;
    const int64_t total = (int64_t)p->dp.w_words + N + H;
    hipLaunchKernelGGL(k_rbm_pack, dim3(grid1d(total, 256)), dim3(256), 0, (hipStream_t)stream,
                       N, H, Np, flat, (uint32_t *)p->dp.rbm_W, (float *)p->dp.rbm_a,
                       (float *)p->dp.rbm_b);
    hipLaunchKernelGGL(k_rbm_colsum, dim3(grid1d(N, 256)), dim3(256), 0, (hipStream_t)stream, N,
                       H, Np, p->dp.rbm_W, (float *)p->dp.rbm_cs);
    if (p->dp.rbh)
        hipLaunchKernelGGL(k_rbm_pack_t, dim3(grid1d((int64_t)N * p->dp.Hp, 256)), dim3(256), 0,
                           (hipStream_t)stream, N, H, p->dp.Hp, flat, (uint16_t *)p->dp.rbm_WT);
    if (p->dp.rbt) {   // new tables; rows carried under the old parameters lapse
        rbt_build_tables(p, (hipStream_t)stream);
        ++p->dp.rbt_gen;
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_orbit_mean(int64_t n, float *v, int32_t n_orbits, const int32_t *offsets,
                    const int32_t *members, void *stream) {
    if (n <= 0 || n > 0x7fffffffLL || n_orbits <= 0)
        return fail(GFMC_EINVAL, "gfmc_orbit_mean: bad size");
    if (!v || !offsets || !members) return fail(GFMC_EINVAL, "gfmc_orbit_mean: NULL buffer");
    hipLaunchKernelGGL(k_orbit_mean, dim3(grid1d(n_orbits, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, n_orbits, offsets, members, v);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int orbit_args(const char *fn, int64_t n, int32_t n_orbits, const void *a, const void *b,
                      const void *c) {
    if (n <= 0 || n > 0x7fffffffLL || n_orbits <= 0)
        return fail(GFMC_EINVAL, std::string(fn) + ": bad size");
    if (!a || !b || !c) return fail(GFMC_EINVAL, std::string(fn) + ": NULL buffer");
    return 0;
}

int gfmc_orbit_sum(int64_t n, const float *v, int32_t n_orbits, const int32_t *offsets,
                   const int32_t *members, float *out, void *stream) {
    int rc = orbit_args("gfmc_orbit_sum", n, n_orbits, v, offsets, members);
    if (rc) return rc;
    if (!out) return fail(GFMC_EINVAL, "gfmc_orbit_sum: NULL output");
    hipLaunchKernelGGL(k_orbit_sum, dim3(grid1d(n_orbits, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, n_orbits, offsets, members, v, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_orbit_axpy(int64_t n, float *v, int32_t n_orbits, const int32_t *offsets,
                    const int32_t *members, const float *x, float alpha, void *stream) {
    int rc = orbit_args("gfmc_orbit_axpy", n, n_orbits, v, offsets, members);
    if (rc) return rc;
    if (!x) return fail(GFMC_EINVAL, "gfmc_orbit_axpy: NULL input");
    hipLaunchKernelGGL(k_orbit_axpy, dim3(grid1d(n_orbits, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, n_orbits, offsets, members, x, alpha, v);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_rbm_orbit_O(const gfmc_problem *p, int64_t n, const uint32_t *spins, int32_t n_orbits,
                     const int32_t *offsets, const int32_t *members, float *O, float *theta,
                     void *stream) {
    int rc = check_rbm("gfmc_rbm_orbit_O", p);
    if (rc) return rc;
    if (n < 1 || n > 65535 || n_orbits < 1)
        return fail(GFMC_EINVAL, "gfmc_rbm_orbit_O: n must be in [1, 65535], n_orbits >= 1");
    if (!spins || !offsets || !members || !O || !theta)
        return fail(GFMC_EINVAL, "gfmc_rbm_orbit_O: NULL buffer");
    rc = gfmc_rbm_theta(p, n, spins, theta, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rbm_orbit_O, dim3(grid1d(n_orbits, 256), (unsigned)n), dim3(256), 0,
                       (hipStream_t)stream, p->dp, n, spins, theta, n_orbits, offsets, members,
                       O);
    HIP_TRY(hipGetLastError());
    return 0;
}

int64_t gfmc_sr_work_doubles(int32_t m) {
    if (m < 1 || m > 8192) return -1;
    return (int64_t)m * m + (5 + SR_MC) * (int64_t)m;
}

int gfmc_sr_solve(int64_t n, int32_t m, const float *O, const float *g, double eps_rel,
                  double eps_abs, int32_t max_iter, double tol, double max_norm,
                  int32_t warm_start, float *delta, double *work, void *stream) {
    if (n < 1 || m < 1 || m > 8192 || max_iter < 0)
        return fail(GFMC_EINVAL, "gfmc_sr_solve: bad size");
    if (!O || !g || !delta || !work) return fail(GFMC_EINVAL, "gfmc_sr_solve: NULL buffer");
    if (!(eps_rel >= 0.0) || !(eps_abs >= 0.0) || !(tol >= 0.0) || !(max_norm >= 0.0))
        return fail(GFMC_EINVAL, "gfmc_sr_solve: eps, tol and max_norm must be >= 0");
    const hipStream_t st = (hipStream_t)stream;
    double *S = work, *mean = work + (int64_t)m * m, *x = mean + m, *r = x + m, *pv = r + m,
           *Ap = pv + m, *part = Ap + m;
    hipLaunchKernelGGL(k_sr_mean_part, dim3(grid1d(m, 256), SR_MC), dim3(256), 0, st, n, m, O,
                       part);
    hipLaunchKernelGGL(k_sr_mean, dim3(grid1d(m, 256)), dim3(256), 0, st, n, m,
                       (const double *)part, mean);
    const unsigned tiles = (unsigned)((m + SR_T - 1) / SR_T);
    hipLaunchKernelGGL(k_sr_cov, dim3(tiles, tiles), dim3(SR_T * SR_T), 0, st, n, m, O, mean, S);
    hipLaunchKernelGGL(k_sr_cg, dim3(1), dim3(1024), 0, st, m, (const double *)S, eps_rel,
                       eps_abs, g, max_iter, tol, max_norm, warm_start ? 1 : 0, x, r, pv, Ap,
                       part, delta);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_rbm_theta(const gfmc_problem *p, int64_t n, const uint32_t *spins, float *theta,
                   void *stream) {
    int rc = check_rbm("gfmc_rbm_theta", p);
    if (rc) return rc;
    if (n < 0) return fail(GFMC_EINVAL, "gfmc_rbm_theta: n < 0");
    if (n == 0) return 0;
    if (!spins || !theta) return fail(GFMC_EINVAL, "gfmc_rbm_theta: NULL buffer");
    const int64_t tiles = (n + 31) / 32 * ((p->dp.H + 31) / 32);
    hipLaunchKernelGGL(k_rbm_theta_mfma, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0,
                       (hipStream_t)stream, p->dp, n, spins, theta);
    HIP_TRY(hipGetLastError());
    return 0;
}

int64_t gfmc_rbm_vmin_work_bytes(const gfmc_problem *p, int64_t n) {
    if (check_rbm("gfmc_rbm_vmin_work_bytes", p) || n < 0) return -1;
    return (int64_t)vmin_layout(p, n).total;
}

int gfmc_rbm_vmin_grad(const gfmc_problem *p, int64_t n, const uint32_t *spins, float clip,
                       float *grad, double *out2, void *work, void *stream) {
    int rc = check_rbm("gfmc_rbm_vmin_grad", p);
    if (rc) return rc;
    if (n < 1 || n > INT32_MAX)
        return fail(GFMC_EINVAL, "gfmc_rbm_vmin_grad: n must be in [1, 2^31)");
    if (!spins || !grad || !out2 || !work)
        return fail(GFMC_EINVAL, "gfmc_rbm_vmin_grad: NULL buffer");
    if (!(clip >= 0.0f)) return fail(GFMC_EINVAL, "gfmc_rbm_vmin_grad: clip must be >= 0");
    const int N = p->dp.N, H = p->dp.H;
    const int vw = vmin_waves(N);
    if (vw == 0)
        return fail(GFMC_EUNSUP, "gfmc_rbm_vmin_grad: the per-site accumulators of 64 hidden "
                                 "units exceed 160 KiB of LDS (N > 312)");
    static bool lds_set = false;
    if (!lds_set) {
        HIP_TRY(hipFuncSetAttribute((const void *)k_vmin_moves,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        lds_set = true;
    }
    const hipStream_t st = (hipStream_t)stream;
    const VminLayout L = vmin_layout(p, n);
    unsigned char *w = (unsigned char *)work;
    double *rates = (double *)(w + L.rates), *e = (double *)(w + L.e_loc);
    float *g = (float *)(w + L.g), *theta = (float *)(w + L.theta), *B = (float *)(w + L.B);
    float *Dpart = (float *)(w + L.Dpart), *bpart = (float *)(w + L.bpart);
    float *apart = (float *)(w + L.apart), *G1part = (float *)(w + L.G1part);
    rc = gfmc_rates(p, n, spins, rates, e, nullptr, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vmin_coef, dim3(1), dim3(RECONF_THREADS), 0, st, n, e, clip, g, out2);
    rc = gfmc_rbm_theta(p, n, spins, theta, stream);
    if (rc) return rc;
    const dim3 grid_m((unsigned)((H + 63) / 64), (unsigned)((L.n_cw + vw - 1) / vw));
    hipLaunchKernelGGL(k_vmin_moves, grid_m, dim3(vw * WAVE), vmin_lds_bytes(N, vw), st, p->dp,
                       n, spins, theta, rates, g, B, Dpart, bpart, apart);
    const int64_t bx_waves = L.n_split * ((H + 31) / 32) * ((N + 31) / 32);
    hipLaunchKernelGGL(k_rbm_bx_mfma, dim3((unsigned)((bx_waves + 3) / 4)), dim3(256), 0, st,
                       p->dp, n, spins, B, G1part);
    const int64_t np = (int64_t)H * N + N + H;
    hipLaunchKernelGGL(k_vmin_reduce, dim3(grid1d(np, 256)), dim3(256), 0, st, N, H, L.n_cw,
                       L.n_split, G1part, Dpart, bpart, apart, grad);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_adam_step(int64_t n, float *param, const float *grad, float *m, float *v, int64_t step,
                   float lr, float beta1, float beta2, float eps, void *stream) {
    if (n < 0 || step < 1) return fail(GFMC_EINVAL, "gfmc_adam_step: n < 0 or step < 1");
    if (n == 0) return 0;
    if (!param || !grad || !m || !v) return fail(GFMC_EINVAL, "gfmc_adam_step: NULL buffer");
    if (!(beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f && eps >= 0.0f))
        return fail(GFMC_EINVAL, "gfmc_adam_step: betas must lie in [0, 1), eps >= 0");
    const double bc1 = 1.0 - std::pow((double)beta1, (double)step);
    const double bc2 = 1.0 - std::pow((double)beta2, (double)step);
    hipLaunchKernelGGL(k_adam, dim3(grid1d(n, 256)), dim3(256), 0, (hipStream_t)stream, n, param,
                       grad, m, v, (float)(lr / bc1), (float)std::sqrt(bc2), beta1, beta2, eps);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
