// gfmc_mlp_train.inl -- local-energy variance minimisation of the MLP guide
// (included at the end of gfmc_hip.hip, after gfmc_mlp.inl).
//
// L = (1/n) sum_i (E_i - mean E)^2 with E_i = diag_i + sum_k r_ik and
// r_ik = bare_k psi(x_i^k) / psi(x_i), so
//   dL/dp = sum_i g_i sum_k r_ik (O(x_i^k) - O(x_i)),  O = d log psi / dp,
// g_i = 2 (E_i - mean E) / n (clamped as for the RBM).  Over the rows of the
// propagation (configuration, then the candidates) this is one backward pass
// of sum_rows c_row log psi(row) with c = -g_i sum_k r_ik on the
// configuration row and g_i r_ik on candidate k.  With straight-through BF16
// rounding of h and the exact slopes s' of the tabulated tanh:
//   dL/db3 = sum c,  dL/dw3_c = sum c s(u_c),
//   G2_c = c w3_c s'(u_c):  dL/db2 = sum G2,  dL/dW2 = G2^T h,
//   G1_k = (G2 W2)_k s'(z_k):  dL/db1 = sum G1,  dL/dW1 = G1^T x.
// Kernels: the propagation's rates and scan, k_mlp_coef (c per row),
// k_mlp_gemm<., true> (the forward pass storing h, the table indices, G2 and
// per-wavefront w3 / b3 partials), k_mlp_gw2 (G2^T [h | 1], split over rows),
// k_mlp_g1 (G2 W2 times s'), k_mlp_gw1 (G1^T [x | 1], split over rows) and
// k_mlp_greduce.  The FP32 operands G2, G1 enter the BF16 MFMAs as hi + lo
// BF16 pairs (2^-16 relative); all sums run in a fixed order.

#define MLP_KS 1024   // rows per split of the split-K products

typedef __attribute__((ext_vector_type(8))) __bf16 mt_bf16x8;
typedef __attribute__((ext_vector_type(16))) float mt_f32x16;

struct MlpTrainWork {
    int64_t rows_cap = 0, n_cap = 0, wave_cap = 0, p2_cap = 0, p1_cap = 0;
    std::vector<void *> allocs, n_allocs;   // row buffers; per-sample buffers
    MlpTws T{};
    float *coef = nullptr, *G1 = nullptr, *p2 = nullptr, *p1 = nullptr, *g = nullptr;
    double *rates = nullptr, *e = nullptr;
};

static void free_mlp_train(MlpTrainWork *t) {
    if (!t) return;
    for (void *a : t->allocs) (void)hipFree(a);
    for (void *a : t->n_allocs) (void)hipFree(a);
    delete t;
}

// c per row: -g_i sum_k r_ik on the configuration row, g_i r_ik on
// candidate k (ascending slots, as the rows), 0 on the padding rows
__global__ __launch_bounds__(256) void k_mlp_coef(DevProblem P, int64_t n, const double *rates,
                                                  const float *g, MlpWs W, float *coef) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t M = P.n_off + 1;
    const int cnt = W.cnt[i];
    const size_t base = (size_t)W.blk_off[i] * MLP_RB;
    const int32_t *cand = W.cand + (size_t)i * P.n_off;
    const double gi = (double)g[i];
    double sum = 0.0;
    for (int c = 0; c < cnt - 1; ++c) {
        const double r = rates[i * M + 1 + cand[c]];
        sum += r;
        coef[base + 1 + c] = (float)(gi * r);
    }
    coef[base] = (float)(-gi * sum);
    const int padded = (cnt + MLP_RB - 1) / MLP_RB * MLP_RB;
    for (int c = cnt; c < padded; ++c) coef[base + c] = 0.0f;
}

// dL/dW2 partials over one row range: wavefront (m, n) tile of 32 x 32 of
// G2^T [h | 1] (m = layer-2 unit, n = layer-1 unit, n = h1: the ones column)
__global__ __launch_bounds__(256) void k_mlp_gw2(int h1, int h2, int64_t rows, int n_tiles,
                                                 MlpTws T, float *part) {
    const int lane = threadIdx.x % WAVE;
    const int64_t wid = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE;
    const int mt = h2 / 32, nt = h1 / 32 + 1;
    if (wid >= (int64_t)n_tiles) return;
    const int tile = (int)(wid % (mt * nt));
    const int64_t split = wid / (mt * nt);
    const int m0 = (tile / nt) * 32, n0 = (tile % nt) * 32;
    const int64_t r_lo = split * MLP_KS, r_hi = r_lo + MLP_KS < rows ? r_lo + MLP_KS : rows;
    const int half = lane >> 5, li = lane & 31;
    mt_f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int64_t r0 = r_lo; r0 < r_hi; r0 += 16) {
        mt_bf16x8 ah, al, bv;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t r = r0 + 8 * half + q;
            const bool ok = r < r_hi;
            const float a = ok ? T.G2[r * h2 + m0 + li] : 0.0f;
            {
                const __bf16 hb = (__bf16)a;
                ah[q] = hb;
                al[q] = (__bf16)(a - (float)hb);
            }
            const int nn = n0 + li;
            uint16_t bb = 0;
            if (ok) bb = nn < h1 ? T.Hb[r * h1 + nn] : (nn == h1 ? (uint16_t)0x3F80 : (uint16_t)0);
            bv[q] = __builtin_bit_cast(__bf16, bb);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bv, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bv, acc, 0, 0, 0);
    }
    const int ldn = h1 + 32;
    float *out = part + split * (int64_t)h2 * ldn;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + 8 * (r / 4) + 4 * half + r % 4;
        out[(int64_t)m * ldn + n0 + li] = acc[r];
    }
}

// G1 = (G2 W2) s'(z): wavefront tile of 32 rows x 32 layer-1 units
__global__ __launch_bounds__(256) void k_mlp_g1(DevProblem P, int64_t rows, MlpTws T,
                                                float *G1) {
    extern __shared__ __align__(16) unsigned char lds_g1[];
    float2 *tab = (float2 *)lds_g1;
    for (int i = threadIdx.x; i < MLP_NT; i += blockDim.x) tab[i] = P.mlp_tab[i];
    __syncthreads();
    const int lane = threadIdx.x % WAVE;
    const int h1 = P.mlp_h1, h2 = P.mlp_h2;
    const int nt = h1 / 32;
    const int64_t wid = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE;
    const int64_t ntile = (rows + 31) / 32 * nt;
    if (wid >= ntile) return;
    const int64_t r0 = (wid / nt) * 32;
    const int k0 = (int)(wid % nt) * 32;
    const int half = lane >> 5, li = lane & 31;
    mt_f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int c0 = 0; c0 < h2; c0 += 16) {
        mt_bf16x8 ah, al, bv;
        const int64_t r = r0 + li;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = c0 + 8 * half + q;
            const float a = r < rows ? T.G2[r * h2 + c] : 0.0f;
            {
                const __bf16 hb = (__bf16)a;
                ah[q] = hb;
                al[q] = (__bf16)(a - (float)hb);
            }
            bv[q] = __builtin_bit_cast(__bf16, P.mlp_W2[(size_t)c * h1 + k0 + li]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bv, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t row = r0 + 8 * (r / 4) + 4 * half + r % 4;
        if (row < rows) {
            const int k = k0 + li;
            G1[row * h1 + k] = acc[r] * mlp_slope(tab, T.Jx[row * h1 + k]);
        }
    }
}

// dL/dW1 partials over one row range: wavefront (m, n) tile of G1^T [x | 1]
// (m = layer-1 unit, n = site, n = N: the ones column); x of a row is its
// walker's configuration with the row's move applied
__global__ __launch_bounds__(256) void k_mlp_gw1(DevProblem P, const uint32_t *spins,
                                                 int64_t rows, int n_tiles, MlpTws T,
                                                 const float *G1, float *part) {
    const int lane = threadIdx.x % WAVE;
    const int h1 = P.mlp_h1, N = P.N;
    const int ldn = (N + 1 + 31) / 32 * 32;
    const int mt = h1 / 32, nt = ldn / 32;
    const int64_t wid = (int64_t)blockIdx.x * 4 + threadIdx.x / WAVE;
    if (wid >= (int64_t)n_tiles) return;
    const int tile = (int)(wid % (mt * nt));
    const int64_t split = wid / (mt * nt);
    const int m0 = (tile / nt) * 32, n0 = (tile % nt) * 32;
    const int64_t r_lo = split * MLP_KS, r_hi = r_lo + MLP_KS < rows ? r_lo + MLP_KS : rows;
    const int half = lane >> 5, li = lane & 31;
    const int nn = n0 + li;
    mt_f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int64_t r0 = r_lo; r0 < r_hi; r0 += 16) {
        mt_bf16x8 ah, al, bv;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t r = r0 + 8 * half + q;
            const bool ok = r < r_hi;
            const float a = ok ? G1[r * h1 + m0 + li] : 0.0f;
            {
                const __bf16 hb = (__bf16)a;
                ah[q] = hb;
                al[q] = (__bf16)(a - (float)hb);
            }
            float x = 0.0f;
            if (ok && nn <= N) {
                if (nn == N) {
                    x = 1.0f;
                } else {
                    const int2 ri = T.rinfo[r];
                    const int si = ri.y & 0xffff, sj = (ri.y >> 16) & 0xffff;
                    const uint32_t *sp = spins + (int64_t)ri.x * P.words;
                    x = ((sp[nn >> 5] >> (nn & 31)) & 1u) ? 1.0f : -1.0f;
                    if (nn == si || nn == sj) x = -x;
                }
            }
            bv[q] = (__bf16)x;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bv, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bv, acc, 0, 0, 0);
    }
    float *out = part + split * (int64_t)h1 * ldn;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + 8 * (r / 4) + 4 * half + r % 4;
        out[(int64_t)m * ldn + n0 + li] = acc[r];
    }
}

// the flat gradient [W1 (h1 x N), b1, W2 (h2 x h1), b2, w3, b3] from the
// partials, each entry summed over the splits / wavefronts in order
__global__ __launch_bounds__(256) void k_mlp_greduce(int N, int h1, int h2, int s1, int s2,
                                                     int n_waves, const float *p1,
                                                     const float *p2, const float *wpart,
                                                     float *grad) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nW1 = (int64_t)h1 * N, nW2 = (int64_t)h2 * h1;
    const int64_t total = nW1 + h1 + nW2 + h2 + h2 + 1;
    if (idx >= total) return;
    const int ld1 = (N + 1 + 31) / 32 * 32, ld2 = h1 + 32;
    float s = 0.0f;
    if (idx < nW1 + h1) {
        const int64_t m = idx < nW1 ? idx / N : idx - nW1;
        const int64_t c = idx < nW1 ? idx % N : N;
        for (int t = 0; t < s1; ++t) s += p1[(int64_t)t * h1 * ld1 + m * ld1 + c];
    } else if (idx < nW1 + h1 + nW2 + h2) {
        const int64_t j = idx - nW1 - h1;
        const int64_t m = j < nW2 ? j / h1 : j - nW2;
        const int64_t c = j < nW2 ? j % h1 : h1;
        for (int t = 0; t < s2; ++t) s += p2[(int64_t)t * h2 * ld2 + m * ld2 + c];
    } else {
        const int64_t c = idx - (nW1 + h1 + nW2 + h2);   // 0..h2: w3, then b3
        for (int t = 0; t < n_waves; ++t) s += wpart[(int64_t)t * (h2 + 1) + c];
    }
    grad[idx] = s;
}

// FP32 parameters -> the problem's tables: W1 and b1 on the layer-1 grid
// (BF16, magnitudes below 2^-9 flushed, above 64 clamped; W1 as integers x
// 2^16), W2 on the BF16 grid (below 2^-24 flushed), b2, w3, b3 in FP32
__global__ __launch_bounds__(256) void k_mlp_pack(int N, int h1, int h2, const float *flat,
                                                  int32_t *W1I, double *b1, uint16_t *W2,
                                                  float *b2, float *w3, float *b3) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nW1 = (int64_t)h1 * N, nW2 = (int64_t)h2 * h1;
    const int64_t total = nW1 + h1 + nW2 + h2 + h2 + 1;
    if (idx >= total) return;
    auto bf = [](float v) -> uint16_t {
        const uint32_t b = __float_as_uint(v);
        uint32_t r = (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
        if (fabsf(__uint_as_float(r << 16)) < 5.9604644775390625e-08f) r = 0;
        return (uint16_t)r;
    };
    auto l1 = [](float v) -> float {
        const uint32_t b = __float_as_uint(v);
        const float r = __uint_as_float(((b + 0x7fffu + ((b >> 16) & 1u)) >> 16) << 16);
        if (fabsf(r) < 0.001953125f) return 0.0f;
        return fabsf(r) > 64.0f ? copysignf(64.0f, r) : r;
    };
    const float v = flat[idx];
    if (idx < nW1) {
        const int64_t k = idx / N, n = idx % N;
        W1I[n * h1 + k] = (int32_t)((double)l1(v) * 65536.0);
    } else if (idx < nW1 + h1) {
        b1[idx - nW1] = (double)l1(v);
    } else if (idx < nW1 + h1 + nW2) {
        W2[idx - nW1 - h1] = bf(v);
    } else if (idx < nW1 + h1 + nW2 + h2) {
        b2[idx - nW1 - h1 - nW2] = v;
    } else if (idx < nW1 + h1 + nW2 + 2 * h2) {
        w3[idx - nW1 - h1 - nW2 - h2] = v;
    } else {
        b3[0] = v;
    }
}

static int check_mlp(const char *fn, const gfmc_problem *p) {
    if (!p) return fail(GFMC_EINVAL, std::string(fn) + ": NULL problem");
    if (p->dp.guide != GFMC_GUIDE_MLP)
        return fail(GFMC_EINVAL, std::string(fn) + ": the problem's guide is not an MLP");
    return 0;
}

static int mt_get(std::vector<void *> &list, size_t bytes, void **dst) {
    void *a = nullptr;
    if (hipMalloc(&a, bytes ? bytes : 16) != hipSuccess)
        return fail(GFMC_ENOMEM, "MLP training: workspace allocation failed");
    list.push_back(a);
    *dst = a;
    return 0;
}

// grow the per-sample buffers (rates, local energies, g) to n samples
static int mlp_train_reserve_n(const gfmc_problem *p, int64_t n) {
    MlpTrainWork *t = p->mlpt;
    if (t->n_cap >= n) return 0;
    for (void *a : t->n_allocs) (void)hipFree(a);
    t->n_allocs.clear();
    t->n_cap = 0;
    int rc = 0;
    void *v = nullptr;
    if (!rc && !(rc = mt_get(t->n_allocs, (size_t)n * (p->dp.n_off + 1) * 8, &v)))
        t->rates = (double *)v;
    if (!rc && !(rc = mt_get(t->n_allocs, (size_t)n * 8, &v))) t->e = (double *)v;
    if (!rc && !(rc = mt_get(t->n_allocs, (size_t)n * 4, &v))) t->g = (float *)v;
    if (rc) {
        for (void *a : t->n_allocs) (void)hipFree(a);
        t->n_allocs.clear();
        return rc;
    }
    t->n_cap = n;
    return 0;
}

// grow the row buffers to rows x (h1, h2) and the partials
static int mlp_train_reserve_rows(const gfmc_problem *p, int64_t rows, int64_t waves,
                                  int64_t p2n, int64_t p1n) {
    MlpTrainWork *t = p->mlpt;
    const DevProblem &P = p->dp;
    if (t->rows_cap >= rows && t->wave_cap >= waves && t->p2_cap >= p2n && t->p1_cap >= p1n)
        return 0;
    for (void *a : t->allocs) (void)hipFree(a);
    t->allocs.clear();
    rows = std::max(rows, t->rows_cap);
    waves = std::max(waves, t->wave_cap);
    p2n = std::max(p2n, t->p2_cap);
    p1n = std::max(p1n, t->p1_cap);
    t->rows_cap = t->wave_cap = t->p2_cap = t->p1_cap = 0;
    const size_t R = (size_t)rows, h1 = P.mlp_h1, h2 = P.mlp_h2;
    std::vector<void *> &L = t->allocs;
    int rc = 0;
    void *v = nullptr;
    if (!rc && !(rc = mt_get(L, R * 4, &v))) {
        t->coef = (float *)v;
        t->T.coef = t->coef;
    }
    if (!rc && !(rc = mt_get(L, R * h1 * 2, &v))) t->T.Hb = (uint16_t *)v;
    if (!rc && !(rc = mt_get(L, R * h1 * 2, &v))) t->T.Jx = (uint16_t *)v;
    if (!rc && !(rc = mt_get(L, R * 8, &v))) t->T.rinfo = (int2 *)v;
    if (!rc && !(rc = mt_get(L, R * h2 * 4, &v))) t->T.G2 = (float *)v;
    if (!rc && !(rc = mt_get(L, R * h1 * 4, &v))) t->G1 = (float *)v;
    if (!rc && !(rc = mt_get(L, (size_t)waves * (h2 + 1) * 4, &v))) t->T.wpart = (float *)v;
    if (!rc && !(rc = mt_get(L, (size_t)p2n * 4, &v))) t->p2 = (float *)v;
    if (!rc && !(rc = mt_get(L, (size_t)p1n * 4, &v))) t->p1 = (float *)v;
    if (rc) {
        for (void *a : L) (void)hipFree(a);
        L.clear();
        return rc;
    }
    t->rows_cap = rows;
    t->wave_cap = waves;
    t->p2_cap = p2n;
    t->p1_cap = p1n;
    return 0;
}

extern "C" {

int64_t gfmc_mlp_n_params(const gfmc_problem *p) {
    if (check_mlp("gfmc_mlp_n_params", p)) return -1;
    const int64_t N = p->dp.N, h1 = p->dp.mlp_h1, h2 = p->dp.mlp_h2;
    return h1 * N + h1 + h2 * h1 + h2 + h2 + 1;
}

int gfmc_mlp_set_params(gfmc_problem *p, const float *flat, void *stream) {
    int rc = check_mlp("gfmc_mlp_set_params", p);
    if (rc) return rc;
    if (!flat) return fail(GFMC_EINVAL, "gfmc_mlp_set_params: NULL parameters");
    const DevProblem &P = p->dp;
    const int64_t total = gfmc_mlp_n_params(p);
    hipLaunchKernelGGL(k_mlp_pack, dim3(grid1d(total, 256)), dim3(256), 0, (hipStream_t)stream,
                       P.N, P.mlp_h1, P.mlp_h2, flat, (int32_t *)P.mlp_W1I, (double *)P.mlp_b1,
                       (uint16_t *)P.mlp_W2, (float *)P.mlp_b2, (float *)P.mlp_w3,
                       (float *)P.mlp_b3p);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_mlp_vmin_grad(const gfmc_problem *p, int64_t n, const uint32_t *spins, float clip,
                       float *grad, double *out2, void *stream) {
    int rc = check_mlp("gfmc_mlp_vmin_grad", p);
    if (rc) return rc;
    if (n < 1 || n > 0x7fffffffLL)
        return fail(GFMC_EINVAL, "gfmc_mlp_vmin_grad: n must be in [1, 2^31)");
    if (!spins || !grad || !out2) return fail(GFMC_EINVAL, "gfmc_mlp_vmin_grad: NULL buffer");
    if (!(clip >= 0.0f)) return fail(GFMC_EINVAL, "gfmc_mlp_vmin_grad: clip must be >= 0");
    if (g_step_base)
        return fail(GFMC_EUNSUP, "gfmc_mlp_vmin_grad synchronises with the host and cannot be "
                                 "captured into a graph");
    const hipStream_t st = (hipStream_t)stream;
    const DevProblem &P = p->dp;
    const int h1 = P.mlp_h1, h2 = P.mlp_h2, N = P.N;
    // the rows (candidate lists, blocks) and the rates of every sample
    rc = mlp_reserve(p, n);
    if (rc) return rc;
    if (!p->mlpt) p->mlpt = new MlpTrainWork;
    rc = mlp_train_reserve_n(p, n);
    if (rc) return rc;
    MlpTrainWork *t = p->mlpt;
    rc = mlp_rates(p, n, spins, t->rates, t->e, nullptr, nullptr, st);
    if (rc) return rc;
    int blocks = 0;
    rc = mlp_read_total(p, st, &blocks);
    if (rc) return rc;
    const int64_t rows = (int64_t)blocks * MLP_RB;
    const int64_t splits = (rows + MLP_KS - 1) / MLP_KS;
    const int ld1 = (N + 1 + 31) / 32 * 32;
    const int64_t p2n = splits * h2 * (int64_t)(h1 + 32), p1n = splits * h1 * (int64_t)ld1;
    const int64_t waves = (int64_t)p->n_cu * 4 * MLP_GW;
    rc = mlp_train_reserve_rows(p, rows, waves, p2n, p1n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vmin_coef, dim3(1), dim3(RECONF_THREADS), 0, st, n, t->e, clip, t->g,
                       out2);
    hipLaunchKernelGGL(k_mlp_coef, dim3(grid1d(n, 256)), dim3(256), 0, st, P, n, t->rates, t->g,
                       p->mlp->ws, t->coef);
    const unsigned fgrid = launch_mlp_gemm<true>(p, spins, st, t->T);
    const int n_waves = (int)fgrid * MLP_GW;
    const int tiles2 = (int)(splits * (h2 / 32) * (h1 / 32 + 1));
    hipLaunchKernelGGL(k_mlp_gw2, dim3((unsigned)((tiles2 + 3) / 4)), dim3(256), 0, st, h1, h2,
                       rows, tiles2, t->T, t->p2);
    const int64_t g1_tiles = (rows + 31) / 32 * (h1 / 32);
    hipLaunchKernelGGL(k_mlp_g1, dim3((unsigned)((g1_tiles + 3) / 4)), dim3(256),
                       (size_t)MLP_NT * sizeof(float2), st, P, rows, t->T, t->G1);
    const int tiles1 = (int)(splits * (h1 / 32) * (ld1 / 32));
    hipLaunchKernelGGL(k_mlp_gw1, dim3((unsigned)((tiles1 + 3) / 4)), dim3(256), 0, st, P, spins,
                       rows, tiles1, t->T, (const float *)t->G1, t->p1);
    const int64_t np = gfmc_mlp_n_params(p);
    hipLaunchKernelGGL(k_mlp_greduce, dim3(grid1d(np, 256)), dim3(256), 0, st, N, h1, h2,
                       (int)splits, (int)splits, n_waves, (const float *)t->p1,
                       (const float *)t->p2, (const float *)t->T.wpart, grad);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
