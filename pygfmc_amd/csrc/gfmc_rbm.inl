// gfmc_rbm.inl -- RBM guide kernels with lanes over hidden units (included by
// gfmc_hip.hip; used when the hidden layer has at most 64 * RBH_MAX_M units).
//
// One wavefront per walker.  Lane l owns the hidden units h = l + 64 m,
// m < M (M = Hp / 64, Hp = H rounded up to a multiple of 64), and keeps
// theta_h, p_h = (1 + tanh theta_h) / 2 and q_h = 1 - p_h in registers.  The
// weights are staged in LDS once per workgroup, transposed and in BF16:
// WT[k][h] with row stride Hp, the padded hidden units zero, so the two weight
// rows a slot needs are read by the wavefront as contiguous 128-byte lines.
// The log-ratio of slot (i, j) is
//   -2 (a_i x_i + a_j x_j) - sum_h Delta_h + sum_h log(p_h e^{2 Delta_h} + q_h),
//   Delta_h = -2 (x_i W_hi + x_j W_hj)
// (cosh(theta + Delta) / cosh theta = e^{-Delta} (p e^{2 Delta} + q)); each
// lane multiplies its M factors, takes one log2, and the wave tree adds the
// 64 partial logarithms; sum_h Delta_h comes from the column sums of W.  A
// padded hidden unit has theta = 0, so p = q = 1/2 and its factor is exactly
// 1.  Walkers are taken from a per-launch queue (an atomic counter that a
// one-thread kernel resets before every propagation launch), so walkers with
// uneven hop counts balance over the persistent grid.

#define RBH_WAVES 12
#define RBH_MAX_M 8

__device__ __forceinline__ float bf16_u16(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

struct RbhLds {
    float *w;            // [K * 64] slot weights (blocked layout, wpos)
    int2 *ent;           // [n_off + 1] active slots: {LDS position | flags, i | j << 16}
    float *g;            // [n_off + 1] their field term + log |bare rate|
    signed char *x;      // [n_pad]
    unsigned char *anti; // [e_pad]
};

// the walker state of wavefront `wave`; `skip` bytes of shared tables come
// first (the BF16 weights of the rbh kernels, the factor tables of the tied
// kernel).  The active-slot list has room for four padding entries.
__device__ __forceinline__ RbhLds rbh_carve_at(const DevProblem &P, unsigned char *lds, int wave,
                                               size_t skip) {
    unsigned char *base = lds + skip + (size_t)wave * P.rbh_wave_bytes;
    RbhLds L;
    L.w = (float *)base;
    L.ent = (int2 *)(L.w + P.K * WAVE);
    L.g = (float *)(L.ent + (P.n_off + 4));
    L.x = (signed char *)(L.g + (P.n_off + 4));
    L.anti = (unsigned char *)(L.x + P.n_pad);
    return L;
}

__device__ __forceinline__ RbhLds rbh_carve(const DevProblem &P, unsigned char *lds, int wave) {
    return rbh_carve_at(P, lds, wave, (size_t)P.wt_bytes);
}

// flags of an active-slot entry (bits above the 24-bit LDS position)
#define RBH_NEG_I (1 << 24)    // x_i = -1
#define RBH_HAS_J (1 << 25)    // two-site move
#define RBH_NEG_J (1 << 26)    // x_j = -1
#define RBH_POS_BARE (1 << 27) // bare rate > 0 (the weight is then negative)

template <int M>
struct RbhState {
    float th[M], p[M], q[M];
};

__device__ void rbh_stage(const DevProblem &P, uint16_t *WT) {
    const uint32_t *src = (const uint32_t *)P.rbm_WT;
    uint32_t *dst = (uint32_t *)WT;
    const int words = P.wt_bytes / 4;
    for (int t = threadIdx.x; t < words; t += blockDim.x) dst[t] = src[t];
    __syncthreads();
}

template <int M>
__device__ __forceinline__ void rbh_pq(RbhState<M> &S) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const float e = __expf(-2.0f * fabsf(S.th[m]));
        const float inv = __builtin_amdgcn_rcpf(1.0f + e);
        const float big = inv, small = e * inv;
        S.p[m] = S.th[m] >= 0.0f ? big : small;
        S.q[m] = S.th[m] >= 0.0f ? small : big;
    }
}

// spins -> x, padded weights zero, theta from scratch (k ascending), p, q
template <int M>
__device__ void rbh_start(const DevProblem &P, const uint16_t *WT, RbhLds &L, RbhState<M> &S,
                          const uint32_t *sp, int lane) {
    for (int k = lane; k < P.N; k += WAVE) L.x[k] = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1 : -1;
    for (int s = P.n_off + lane; s < P.K * WAVE; s += WAVE) L.w[wpos(s, P.K)] = 0.0f;
    wave_sync();
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int h = lane + 64 * m;
        S.th[m] = h < P.H ? P.rbm_b[h] : 0.0f;
    }
    const uint16_t *col = WT + lane;
    for (int k = 0; k < P.N; ++k) {
        const float xk = (float)L.x[k];
#pragma unroll
        for (int m = 0; m < M; ++m) S.th[m] += bf16_u16(col[(size_t)k * P.Hp + 64 * m]) * xk;
    }
    rbh_pq(S);
}

// this lane's share of log2 prod_h (p_h e^{2 Delta_h} + q_h) for flipping i
// (and j >= 0); ci = -4 log2(e) x_i, cj likewise
template <int M>
__device__ __forceinline__ float rbh_part(const DevProblem &P, const uint16_t *WT,
                                          const RbhState<M> &S, int i, int j, float ci, float cj,
                                          int lane) {
    const uint16_t *ri = WT + (size_t)i * P.Hp + lane;
    const uint16_t *rj = WT + (size_t)(j >= 0 ? j : i) * P.Hp + lane;
    const float cjj = j >= 0 ? cj : 0.0f;
    float a[M];
#pragma unroll
    for (int m = 0; m < M; ++m) a[m] = fmaf(cjj, bf16_u16(rj[64 * m]), ci * bf16_u16(ri[64 * m]));
    float prod = 1.0f;
#pragma unroll
    for (int m = 0; m < M; ++m) prod *= fmaf(S.p[m], __builtin_amdgcn_exp2f(a[m]), S.q[m]);
    if (prod > 1e-30f && prod < 1e30f) return __builtin_amdgcn_logf(prod);
    float acc = 0.0f;   // rare: factors far from 1, one logarithm each
#pragma unroll
    for (int m = 0; m < M; ++m)
        acc += __builtin_amdgcn_logf(fmaf(S.p[m], __builtin_amdgcn_exp2f(a[m]), S.q[m]));
    return acc;
}

// this lane's shares of log2 prod_h (p_h e^{2 Delta_h} + q_h) for two list
// entries at once (independent chains for the scheduler); the second row is
// i again with cj = 0 for a site flip.  A product outside [1e-30, 1e30] in
// any lane (factors far from 1) sends the wavefront to one logarithm per
// factor.
template <int M>
__device__ __forceinline__ void rbh_part_pair(const DevProblem &P, const uint16_t *WT,
                                              const RbhState<M> &S, int i0, int j0, float ci0,
                                              float cj0, int i1, int j1, float ci1, float cj1,
                                              int lane, float *out0, float *out1) {
    const uint16_t *ri0 = WT + (size_t)i0 * P.Hp + lane, *rj0 = WT + (size_t)j0 * P.Hp + lane;
    const uint16_t *ri1 = WT + (size_t)i1 * P.Hp + lane, *rj1 = WT + (size_t)j1 * P.Hp + lane;
    float prod0 = 1.0f, prod1 = 1.0f;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const float a0 = fmaf(cj0, bf16_u16(rj0[64 * m]), ci0 * bf16_u16(ri0[64 * m]));
        const float a1 = fmaf(cj1, bf16_u16(rj1[64 * m]), ci1 * bf16_u16(ri1[64 * m]));
        prod0 *= fmaf(S.p[m], __builtin_amdgcn_exp2f(a0), S.q[m]);
        prod1 *= fmaf(S.p[m], __builtin_amdgcn_exp2f(a1), S.q[m]);
    }
    const bool ok = prod0 > 1e-30f && prod0 < 1e30f && prod1 > 1e-30f && prod1 < 1e30f;
    if (__ballot(!ok) == 0ull) {
        *out0 = __builtin_amdgcn_logf(prod0);
        *out1 = __builtin_amdgcn_logf(prod1);
        return;
    }
    float acc0 = 0.0f, acc1 = 0.0f;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const float a0 = fmaf(cj0, bf16_u16(rj0[64 * m]), ci0 * bf16_u16(ri0[64 * m]));
        const float a1 = fmaf(cj1, bf16_u16(rj1[64 * m]), ci1 * bf16_u16(ri1[64 * m]));
        acc0 += __builtin_amdgcn_logf(fmaf(S.p[m], __builtin_amdgcn_exp2f(a0), S.q[m]));
        acc1 += __builtin_amdgcn_logf(fmaf(S.p[m], __builtin_amdgcn_exp2f(a1), S.q[m]));
    }
    *out0 = acc0;
    *out1 = acc1;
}

__device__ __forceinline__ float rbh_field(const DevProblem &P, int i, int j, float xi, float xj) {
    float sum_delta = -2.0f * xi * P.rbm_cs[i];
    float afield = -2.0f * P.rbm_a[i] * xi;
    if (j >= 0) {
        sum_delta += -2.0f * xj * P.rbm_cs[j];
        afield += -2.0f * P.rbm_a[j] * xj;
    }
    return afield - sum_delta;
}

// log psi(x') - log psi(x) of slot s (wave-uniform s)
template <int M>
__device__ __forceinline__ float rbh_log_ratio(const DevProblem &P, const uint16_t *WT,
                                               const RbhLds &L, const RbhState<M> &S, int s,
                                               int lane) {
    int i, j;
    slot_sites(P, s, &i, &j);
    const float L2E = 1.4426950408889634f;
    const float xi = (float)L.x[i], xj = j >= 0 ? (float)L.x[j] : 0.0f;
    const float part = rbh_part<M>(P, WT, S, i, j, -4.0f * L2E * xi, -4.0f * L2E * xj, lane);
    const float tot = lane_bcast(wave_incl_scan(part, lane), WAVE - 1);
    return rbh_field(P, i, j, xi, xj) + 0.69314718055994531f * tot;
}

__device__ __forceinline__ float rbh_bare(const DevProblem &P, const Tables<float> &tb,
                                          const RbhLds &L, int s) {
    if (s < P.n_hx) return tb.hx_half[s];
    const int pk = P.edge_pack[s - P.n_hx];
    return (L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff]) ? tb.jpm[s - P.n_hx]
                                                          : tb.jpmpm[s - P.n_hx];
}

// every slot weight and the edge states; returns the anti count.  The slots
// with a non-zero bare rate are listed first (lanes over slots: LDS position,
// sites, signs, field term + log |bare|), then evaluated two at a time by
// the whole wavefront (an odd count repeats the last entry).
template <int M>
__device__ int rbh_eval_all(const DevProblem &P, const Tables<float> &tb, const uint16_t *WT,
                            RbhLds &L, const RbhState<M> &S, int lane) {
    int n_act = 0;
    for (int c = 0; c < P.n_off; c += WAVE) {
        const int s = c + lane;
        float bare = 0.0f;
        int i = 0, j = -1;
        if (s < P.n_off) {
            slot_sites(P, s, &i, &j);
            bare = s < P.n_hx ? tb.hx_half[s]
                              : ((L.x[i] != L.x[j]) ? tb.jpm[s - P.n_hx] : tb.jpmpm[s - P.n_hx]);
        }
        const bool act = s < P.n_off && bare != 0.0f;
        if (s < P.n_off && !act) L.w[wpos(s, P.K)] = -bare;
        const unsigned long long msk = __ballot(act);
        const int pos = n_act + __popcll(msk & ((1ull << lane) - 1ull));
        if (act) {
            const float xi = (float)L.x[i], xj = j >= 0 ? (float)L.x[j] : 0.0f;
            int a = wpos(s, P.K);
            if (xi < 0.0f) a |= RBH_NEG_I;
            if (j >= 0) a |= RBH_HAS_J;
            if (xj < 0.0f) a |= RBH_NEG_J;
            if (bare > 0.0f) a |= RBH_POS_BARE;
            L.ent[pos] = make_int2(a, i | ((j >= 0 ? j : i) << 16));
            L.g[pos] = rbh_field(P, i, j, xi, xj) + __logf(fabsf(bare));
        }
        n_act += __popcll(msk);
    }
    wave_sync();
    if ((n_act & 1) && lane == 0) {   // pad to an even count: repeat the last entry
        L.ent[n_act] = L.ent[n_act - 1];
        L.g[n_act] = L.g[n_act - 1];
    }
    wave_sync();
    const float C = 4.0f * 1.4426950408889634f;
    for (int t = 0; t < n_act; t += 2) {
        const int4 e2 = *(const int4 *)(L.ent + t);
        const float2 g2 = *(const float2 *)(L.g + t);
        const int a0 = __builtin_amdgcn_readfirstlane(e2.x), b0 = __builtin_amdgcn_readfirstlane(e2.y);
        const int a1 = __builtin_amdgcn_readfirstlane(e2.z), b1 = __builtin_amdgcn_readfirstlane(e2.w);
        const float ci0 = (a0 & RBH_NEG_I) ? C : -C;
        const float cj0 = (a0 & RBH_HAS_J) ? ((a0 & RBH_NEG_J) ? C : -C) : 0.0f;
        const float ci1 = (a1 & RBH_NEG_I) ? C : -C;
        const float cj1 = (a1 & RBH_HAS_J) ? ((a1 & RBH_NEG_J) ? C : -C) : 0.0f;
        float pa, pb;
        rbh_part_pair<M>(P, WT, S, b0 & 0xffff, (b0 >> 16) & 0xffff, ci0, cj0, b1 & 0xffff,
                         (b1 >> 16) & 0xffff, ci1, cj1, lane, &pa, &pb);
        const float ta = lane_bcast(wave_incl_scan(pa, lane), WAVE - 1);
        const float tbv = lane_bcast(wave_incl_scan(pb, lane), WAVE - 1);
        if (lane == 0) {
            const float wa = __expf(g2.x + 0.69314718055994531f * ta);
            const float wb = __expf(g2.y + 0.69314718055994531f * tbv);
            L.w[a0 & 0xffffff] = (a0 & RBH_POS_BARE) ? -wa : wa;
            L.w[a1 & 0xffffff] = (a1 & RBH_POS_BARE) ? -wb : wb;
        }
    }
    int cnt = 0;
    if (P.has_jz) {
        for (int e = lane; e < P.n_edges; e += WAVE) {
            const int pk = P.edge_pack[e];
            const unsigned char a = L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff];
            if (P.flags) L.anti[e] = a;
            cnt += a;
        }
    }
    wave_sync();
    return wave_sum_int(cnt);
}

// theta += Delta of flipping i (and j >= 0) at the current (pre-flip) x
template <int M>
__device__ __forceinline__ void rbh_move_theta(const DevProblem &P, const uint16_t *WT,
                                               const RbhLds &L, RbhState<M> &S, int i, int j,
                                               int lane) {
    const float xi = (float)L.x[i], xj = j >= 0 ? (float)L.x[j] : 0.0f;
    const uint16_t *ri = WT + (size_t)i * P.Hp + lane;
    const uint16_t *rj = WT + (size_t)(j >= 0 ? j : i) * P.Hp + lane;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        float d = -2.0f * xi * bf16_u16(ri[64 * m]);
        if (j >= 0) d += -2.0f * xj * bf16_u16(rj[64 * m]);
        S.th[m] = S.th[m] + d;
    }
    rbh_pq(S);
}

__global__ void k_rbh_queue_reset(int *queue) {
    if (threadIdx.x == 0) *queue = 0;
}

// the next walker of this launch's queue (wave-uniform)
__device__ __forceinline__ int64_t rbh_next(int *queue, int lane) {
    int v = 0;
    if (lane == 0) v = atomicAdd(queue, 1);
    return (int64_t)__builtin_amdgcn_readfirstlane(v);
}

template <int M>
__global__ __launch_bounds__(RBH_WAVES * WAVE) void k_propagate_rbh(
    DevProblem P, Tables<float> tb, int64_t nw, uint32_t *spins, double *log_w, double *e_loc,
    double beta, double alw, uint32_t key0, uint32_t key1, uint32_t step, uint32_t tag,
    int32_t *evals, const uint32_t *step_base) {
    step += step_base ? *step_base : 0u;
    extern __shared__ __align__(16) unsigned char lds[];
    uint16_t *WT = (uint16_t *)lds;
    rbh_stage(P, WT);
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    RbhLds L = rbh_carve(P, lds, wave);
    WalkerLds<float> WL;
    WL.w = L.w;
    WL.x = L.x;
    WL.anti = L.anti;
    RbhState<M> S;
    for (int64_t wk = rbh_next(P.rbm_queue, lane); wk < nw; wk = rbh_next(P.rbm_queue, lane)) {
        uint32_t *sp = spins + wk * P.words;
        rbh_start<M>(P, WT, L, S, sp, lane);
        int n_anti = rbh_eval_all<M>(P, tb, WT, L, S, lane);
        double diag = diag_value(P, L.x, L.anti, n_anti, lane);
        double lw = log_w[wk];
        double t = 0.0, eL = 0.0;
        uint32_t it = 0;
        while (t < beta) {
            float lane_sum;
            const float incl = lane_scan(P, WL, lane, &lane_sum);
            const float R = lane_bcast(incl, WAVE - 1);
            eL = diag - (double)R;
            const gfmc_u4 r = gfmc_philox(it, (uint32_t)wk, step, tag, key0, key1);
            const double psi = gfmc_u53(r.z, r.w);
            int idx = -1;
            double tau_raw = INFINITY;
            if (R > 0.0f) {
                float u = gfmc_u24(r.x) * R;
                const float cap = below(R);
                if (u > cap) u = cap;
                idx = sample_slot(P, WL, incl, lane_sum, u, lane);
                tau_raw = gfmc_neg_log_unit(1.0 - psi) / (double)R;
            }
            const double rem = beta - t;
            const double tau = tau_raw < rem ? tau_raw : rem;
            lw = lw - eL * tau;
            t = t + tau;
            ++it;
            if (tau_raw < rem && idx >= 0) {
                int i, j;
                slot_sites(P, idx, &i, &j);
                rbh_move_theta<M>(P, WT, L, S, i, j, lane);
                wave_sync();
                if (lane == 0) {
                    L.x[i] = (signed char)-L.x[i];
                    if (j >= 0) L.x[j] = (signed char)-L.x[j];
                }
                wave_sync();
                n_anti = rbh_eval_all<M>(P, tb, WT, L, S, lane);
                diag = diag_value(P, L.x, L.anti, n_anti, lane);
            }
        }
        lw = lw - alw * beta;
        store_spins(P, sp, L.x, lane);
        if (lane == 0) {
            log_w[wk] = lw;
            e_loc[wk] = eL;
            if (evals) evals[wk] += (int32_t)it;
        }
        wave_sync();
    }
}

template <int M>
__global__ __launch_bounds__(RBH_WAVES * WAVE) void k_rates_rbh(
    DevProblem P, Tables<float> tb, int64_t nw, const uint32_t *spins, double *rates,
    double *e_loc, double *log_diff, double *log_psi) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint16_t *WT = (uint16_t *)lds;
    rbh_stage(P, WT);
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    RbhLds L = rbh_carve(P, lds, wave);
    WalkerLds<float> WL;
    WL.w = L.w;
    WL.x = L.x;
    WL.anti = L.anti;
    RbhState<M> S;
    const int64_t Mt = P.n_off + 1;
    for (int64_t wk = (int64_t)blockIdx.x * RBH_WAVES + wave; wk < nw;
         wk += (int64_t)gridDim.x * RBH_WAVES) {
        rbh_start<M>(P, WT, L, S, spins + wk * P.words, lane);
        if (log_psi) {
            float acc = 0.0f;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const float a = fabsf(S.th[m]);
                if (lane + 64 * m < P.H) acc += a + log1pf(__expf(-2.0f * a)) - 0.69314718055994531f;
            }
            for (int k = lane; k < P.N; k += WAVE) acc += P.rbm_a[k] * (float)L.x[k];
            const float tot = lane_bcast(wave_incl_scan(acc, lane), WAVE - 1);
            if (lane == 0) log_psi[wk] = (double)tot;
        }
        if (rates || e_loc || log_diff) {
            const int n_anti = rbh_eval_all<M>(P, tb, WT, L, S, lane);
            const double diag = diag_value(P, L.x, L.anti, n_anti, lane);
            float lane_sum;
            const float incl = lane_scan(P, WL, lane, &lane_sum);
            const float R = lane_bcast(incl, WAVE - 1);
            if (rates) {
                if (lane == 0) rates[wk * Mt] = diag;
                for (int s = lane; s < P.n_off; s += WAVE)
                    rates[wk * Mt + 1 + s] = -(double)L.w[wpos(s, P.K)];
            }
            if (log_diff) {
                if (lane == 0) log_diff[wk * Mt] = 0.0;
                for (int s = 0; s < P.n_off; ++s) {
                    const float lr = rbh_log_ratio<M>(P, WT, L, S, s, lane);
                    if (lane == 0) log_diff[wk * Mt + 1 + s] = (double)lr;
                }
            }
            if (e_loc && lane == 0) e_loc[wk] = diag - (double)R;
        }
        wave_sync();
    }
}

// FP32 parameters -> the transposed BF16 rows WT[k][h] (padded units zero)
__global__ __launch_bounds__(256) void k_rbm_pack_t(int N, int H, int Hp, const float *flat,
                                                    uint16_t *WT) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)N * Hp) return;
    const int k = (int)(q / Hp), h = (int)(q % Hp);
    if (h >= H) {
        WT[q] = 0;
        return;
    }
    const uint32_t b = __float_as_uint(flat[(int64_t)h * N + k]);
    WT[q] = (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);   // nearest even
}

// ---------------------------------------------------------------------------
// translation-tied RBM on a torus (gfmc_rbm_set_tied)
// ---------------------------------------------------------------------------
// Hidden unit h = f n_T + r sees W[h][k] = w_f(k - r).  The factor of unit h
// in the ratio of pair slot (i, j) with x_j = -x_i,
//   p_h 2^{c_i W_hi + c_j W_hj} + q_h,  c_i = -4 log2(e) x_i,  c_j = -c_i,
// depends on the walker only through p_h, q_h: the power is a function of
// (slot class, x_i, f, the site i - r) and sits in a table,
//   tab[((cls 2 + (x_i < 0)) F + f) Q + q],  q = code(i) + off(h) - f Q,
// computed with the expression the rbh kernels evaluate per factor (the
// same bits).  Block 2 n_cls holds w_f itself (theta updates), then Q
// zeros that the padded hidden units read.  Per lane and chunk m the kernel
// keeps off(h) in a register; a slot adds its block base and code(i) (one
// integer add, one LDS read, one fma and one multiply per factor).  Slots are
// evaluated four at a time and their lane products combined together:
// v_permlane32_swap pairs lanes l and l + 32 of two slots at once,
// v_permlane16_swap rows 0 / 1 and 2 / 3 of the two pairs (products, so each
// lane then holds 4 M factors of one slot: one log2 per lane for four slots),
// then a row scan (row_shr 1, 2, 4, 8) of the logarithms leaves slot
// t + {0, 2, 1, 3} in lane 15 of rows 0..3.  Each slot's value is
// sum_{l<16} log2((x_l x_{l+32}) (x_{l+16} x_{l+48})) in the row scan's
// order: the same tree for every slot, so a slot's weight does not depend on
// its place in the list (unless the quad takes the per-factor fallback).

__global__ __launch_bounds__(256) void k_rbt_tables(DevProblem P, float *tab) {
    const int F = P.rbt_F, Q = P.rbt_Q, nc = P.rbt_ncls;
    const int64_t fq = (int64_t)F * Q;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)(2 * nc + 1) * fq + Q;
    if (t >= total) return;
    const int64_t blk = t / fq;
    if (blk > 2 * nc) {   // the zero block
        tab[t] = 0.0f;
        return;
    }
    const int q = (int)(t % Q), f = (int)((t / Q) % F);
    const int hrow = f * P.rbt_nT;   // r = 0: the identity translation
    const float wi = bf16_u16(P.rbm_WT[(size_t)P.rbt_qsite[q] * P.Hp + hrow]);
    if (blk == 2 * nc) {
        tab[t] = wi;
        return;
    }
    const int c = (int)(blk >> 1), sg = (int)(blk & 1);
    const int d2 = P.rbt_qpair[(size_t)c * Q + q];
    if (d2 < 0) {   // an entry of another basis site: never read
        tab[t] = 1.0f;
        return;
    }
    const float wj = bf16_u16(P.rbm_WT[(size_t)d2 * P.Hp + hrow]);
    const float C = 4.0f * 1.4426950408889634f;
    const float ci = sg ? C : -C, cj = -ci;
    tab[t] = __builtin_amdgcn_exp2f(fmaf(cj, wj, ci * wi));
}

// p, q of the lane's units; padded units (h >= H, last chunk only) get p = 0,
// q = 1: their factor is exactly 1
template <int M>
__device__ __forceinline__ void rbt_pq(RbhState<M> &S, bool pad_last) {
    rbh_pq(S);
    if (pad_last) {
        S.p[M - 1] = 0.0f;
        S.q[M - 1] = 1.0f;
    }
}

template <int M>
__device__ void rbt_start(const DevProblem &P, const float *Wt, RbhLds &L, RbhState<M> &S,
                          const int *vb, bool pad_last, const uint32_t *sp, int lane) {
    for (int k = lane; k < P.N; k += WAVE) L.x[k] = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1 : -1;
    for (int s = P.n_off + lane; s < P.K * WAVE; s += WAVE) L.w[wpos(s, P.K)] = 0.0f;
    wave_sync();
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int h = lane + 64 * m;
        S.th[m] = h < P.H ? P.rbm_b[h] : 0.0f;
    }
    for (int k = 0; k < P.N; ++k) {
        const float xk = (float)L.x[k];
        const int ck = P.rbt_kcode[k];
#pragma unroll
        for (int m = 0; m < M; ++m) S.th[m] += Wt[ck + vb[m]] * xk;
    }
    rbt_pq(S, pad_last);
}

template <int M>
__device__ __forceinline__ void rbt_move_theta(const DevProblem &P, const float *Wt,
                                               const RbhLds &L, RbhState<M> &S, const int *vb,
                                               bool pad_last, int i, int j) {
    const float xi = (float)L.x[i], xj = (float)L.x[j];
    const int ci = P.rbt_kcode[i], cj = P.rbt_kcode[j];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        float d = -2.0f * xi * Wt[ci + vb[m]];
        d += -2.0f * xj * Wt[cj + vb[m]];
        S.th[m] = S.th[m] + d;
    }
    rbt_pq(S, pad_last);
}

__device__ __forceinline__ float as_f(uint32_t v) { return __uint_as_float(v); }
__device__ __forceinline__ uint32_t as_u(float v) { return __float_as_uint(v); }

// every slot weight and the edge states; returns the anti count (the rbh
// kernels' rbh_eval_all with table factors, four slots at a time)
template <int M>
__device__ int rbt_eval_all(const DevProblem &P, const Tables<float> &tb, const float *T,
                            RbhLds &L, const RbhState<M> &S, const int *vb, int lane) {
    const int FQ = P.rbt_F * P.rbt_Q;
    int n_act = 0;
    for (int c = 0; c < P.n_off; c += WAVE) {
        const int s = c + lane;
        float bare = 0.0f;
        int i = 0, j = 0;
        if (s < P.n_off) {
            slot_sites(P, s, &i, &j);
            bare = (L.x[i] != L.x[j]) ? tb.jpm[s] : tb.jpmpm[s];
        }
        const bool act = s < P.n_off && bare != 0.0f;
        if (s < P.n_off && !act) L.w[wpos(s, P.K)] = -bare;
        const unsigned long long msk = __ballot(act);
        const int pos = n_act + __popcll(msk & ((1ull << lane) - 1ull));
        if (act) {
            const float xi = (float)L.x[i], xj = (float)L.x[j];
            int a = wpos(s, P.K);
            if (bare > 0.0f) a |= RBH_POS_BARE;
            const int sc = P.rbt_scode[s];
            L.ent[pos] = make_int2(a, ((sc >> 20) * 2 + (xi < 0.0f ? 1 : 0)) * FQ + (sc & 0xfffff));
            L.g[pos] = rbh_field(P, i, j, xi, xj) + __logf(fabsf(bare));
        }
        n_act += __popcll(msk);
    }
    wave_sync();
    const int padn = (4 - (n_act & 3)) & 3;
    if (n_act > 0 && lane < padn) {   // a multiple of four: repeat the last entry
        L.ent[n_act + lane] = L.ent[n_act - 1];
        L.g[n_act + lane] = L.g[n_act - 1];
    }
    wave_sync();
    const int row = lane >> 4;
    const int mine = row == 0 ? 0 : (row == 1 ? 2 : (row == 2 ? 1 : 3));
    for (int t = 0; t < n_act; t += 4) {
        const int4 e01 = *(const int4 *)(L.ent + t);
        const int4 e23 = *(const int4 *)(L.ent + t + 2);
        const int b0 = __builtin_amdgcn_readfirstlane(e01.y);
        const int b1 = __builtin_amdgcn_readfirstlane(e01.w);
        const int b2 = __builtin_amdgcn_readfirstlane(e23.y);
        const int b3 = __builtin_amdgcn_readfirstlane(e23.w);
        float p0 = 1.0f, p1 = 1.0f, p2 = 1.0f, p3 = 1.0f;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const float g0 = T[b0 + vb[m]], g1 = T[b1 + vb[m]];
            const float g2 = T[b2 + vb[m]], g3 = T[b3 + vb[m]];
            p0 *= fmaf(S.p[m], g0, S.q[m]);
            p1 *= fmaf(S.p[m], g1, S.q[m]);
            p2 *= fmaf(S.p[m], g2, S.q[m]);
            p3 *= fmaf(S.p[m], g3, S.q[m]);
        }
        // the four lane products paired across lanes first (l with l + 32,
        // then with l + 16: 28 factors per lane, one slot per row), one
        // logarithm per lane, the row scan; a product outside [1e-30, 1e30]
        // in any lane sends the wavefront to one logarithm per factor, summed
        // in the same tree
        const auto q01 = __builtin_amdgcn_permlane32_swap(as_u(p0), as_u(p1), false, false);
        const auto q23 = __builtin_amdgcn_permlane32_swap(as_u(p2), as_u(p3), false, false);
        const float m01 = as_f(q01[0]) * as_f(q01[1]);
        const float m23 = as_f(q23[0]) * as_f(q23[1]);
        const auto qq = __builtin_amdgcn_permlane16_swap(as_u(m01), as_u(m23), false, false);
        const float pr = as_f(qq[0]) * as_f(qq[1]);
        float v;
        if (__ballot(!(pr > 1e-30f && pr < 1e30f)) == 0ull) {
            v = __builtin_amdgcn_logf(pr);
        } else {   // rare: factors far from 1
            float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f, l3 = 0.0f;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                l0 += __builtin_amdgcn_logf(fmaf(S.p[m], T[b0 + vb[m]], S.q[m]));
                l1 += __builtin_amdgcn_logf(fmaf(S.p[m], T[b1 + vb[m]], S.q[m]));
                l2 += __builtin_amdgcn_logf(fmaf(S.p[m], T[b2 + vb[m]], S.q[m]));
                l3 += __builtin_amdgcn_logf(fmaf(S.p[m], T[b3 + vb[m]], S.q[m]));
            }
            const auto r01 = __builtin_amdgcn_permlane32_swap(as_u(l0), as_u(l1), false, false);
            const auto r23 = __builtin_amdgcn_permlane32_swap(as_u(l2), as_u(l3), false, false);
            const float s01 = as_f(r01[0]) + as_f(r01[1]);
            const float s23 = as_f(r23[0]) + as_f(r23[1]);
            const auto r = __builtin_amdgcn_permlane16_swap(as_u(s01), as_u(s23), false, false);
            v = as_f(r[0]) + as_f(r[1]);
        }
        // rows 0..3: slots t, t + 2, t + 1, t + 3
        v += dpp_get<0x111, 0xf>(v);
        v += dpp_get<0x112, 0xf>(v);
        v += dpp_get<0x114, 0xf>(v);
        v += dpp_get<0x118, 0xf>(v);
        if ((lane & 15) == 15) {
            const int a = L.ent[t + mine].x;
            const float w = __expf(L.g[t + mine] + 0.69314718055994531f * v);
            L.w[a & 0xffffff] = (a & RBH_POS_BARE) ? -w : w;
        }
    }
    int cnt = 0;
    if (P.has_jz) {
        for (int e = lane; e < P.n_edges; e += WAVE) {
            const int pk = P.edge_pack[e];
            const unsigned char a = L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff];
            if (P.flags) L.anti[e] = a;
            cnt += a;
        }
    }
    wave_sync();
    return wave_sum_int(cnt);
}

// the continuous-time propagation of k_propagate_rbh with table factors;
// the factor tables are staged in LDS once per workgroup, walkers come from
// the same per-launch queue
template <int M>
__global__ __launch_bounds__(1024) void k_propagate_rbt(
    DevProblem P, Tables<float> tb, int64_t nw, uint32_t *spins, double *log_w, double *e_loc,
    double beta, double alw, uint32_t key0, uint32_t key1, uint32_t step, uint32_t tag,
    int32_t *evals, const uint32_t *step_base) {
    step += step_base ? *step_base : 0u;
    extern __shared__ __align__(16) unsigned char lds[];
    float *T = (float *)lds;
    {
        const uint4 *src = (const uint4 *)P.rbt_tab;
        uint4 *dst = (uint4 *)T;
        for (int t = threadIdx.x; t < P.rbt_words / 4; t += blockDim.x) dst[t] = src[t];
        __syncthreads();
    }
    const float *Wt = T + (size_t)2 * P.rbt_ncls * P.rbt_F * P.rbt_Q;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    RbhLds L = rbh_carve_at(P, lds, wave, (size_t)P.rbt_words * 4);
    WalkerLds<float> WL;
    WL.w = L.w;
    WL.x = L.x;
    WL.anti = L.anti;
    int vb[M];
#pragma unroll
    for (int m = 0; m < M; ++m) vb[m] = P.rbt_unit[lane + 64 * m];
    const bool pad_last = lane + 64 * (M - 1) >= P.H;
    RbhState<M> S;
    for (int64_t wk = rbh_next(P.rbm_queue, lane); wk < nw; wk = rbh_next(P.rbm_queue, lane)) {
        uint32_t *sp = spins + wk * P.words;
        rbt_start<M>(P, Wt, L, S, vb, pad_last, sp, lane);
        int n_anti = rbt_eval_all<M>(P, tb, T, L, S, vb, lane);
        double diag = diag_value(P, L.x, L.anti, n_anti, lane);
        double lw = log_w[wk];
        double t = 0.0, eL = 0.0;
        uint32_t it = 0;
        while (t < beta) {
            float lane_sum;
            const float incl = lane_scan(P, WL, lane, &lane_sum);
            const float R = lane_bcast(incl, WAVE - 1);
            eL = diag - (double)R;
            const gfmc_u4 r = gfmc_philox(it, (uint32_t)wk, step, tag, key0, key1);
            const double psi = gfmc_u53(r.z, r.w);
            int idx = -1;
            double tau_raw = INFINITY;
            if (R > 0.0f) {
                float u = gfmc_u24(r.x) * R;
                const float cap = below(R);
                if (u > cap) u = cap;
                idx = sample_slot(P, WL, incl, lane_sum, u, lane);
                tau_raw = gfmc_neg_log_unit(1.0 - psi) / (double)R;
            }
            const double rem = beta - t;
            const double tau = tau_raw < rem ? tau_raw : rem;
            lw = lw - eL * tau;
            t = t + tau;
            ++it;
            if (tau_raw < rem && idx >= 0) {
                int i, j;
                slot_sites(P, idx, &i, &j);
                rbt_move_theta<M>(P, Wt, L, S, vb, pad_last, i, j);
                wave_sync();
                if (lane == 0) {
                    L.x[i] = (signed char)-L.x[i];
                    L.x[j] = (signed char)-L.x[j];
                }
                wave_sync();
                n_anti = rbt_eval_all<M>(P, tb, T, L, S, vb, lane);
                diag = diag_value(P, L.x, L.anti, n_anti, lane);
            }
        }
        lw = lw - alw * beta;
        store_spins(P, sp, L.x, lane);
        if (lane == 0) {
            log_w[wk] = lw;
            e_loc[wk] = eL;
            if (evals) evals[wk] += (int32_t)it;
        }
        wave_sync();
    }
}
