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
            if (hop_rate(R)) {
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
// (slot class, x_i, f, the site i - r) and sits in a table T,
//   T[((cls 2 + (x_i < 0)) F + f) Q + q],  q = code(i) + off(h) - f Q,
// computed with the expression the rbh kernels evaluate per factor (the
// same bits); a slot is read from its canonical end i (its class and the
// reverse class share a block).  Block 2 n_cls holds w_f itself (theta
// updates), then Q zeros.  Hidden units come in pairs (f, r), (f, r + d)
// with off(b) = off(a) - D, D = code(d): entry idx of the LDS table is the
// float2 (T[idx], T[idx - D]), so one 8-byte LDS read gives both factors of
// a pair.  A lane owns CP pair chunks (registers 2c, 2c + 1) and CS single
// chunks (register 2 CP + c, the .x half); per chunk it keeps the byte offset
// 8 off(a) of its unit, a slot adds its byte base (one integer add, one LDS
// read, then one fma and one multiply per factor).  Padded registers get
// p = 0, q = 1: their factor is exactly 1.  A single chunk of at most 16
// units is packed (PK): its units sit in lanes l % 16 of every row, and each
// row evaluates it for the slot that row ends up holding, after the lane
// products are combined.  Slots are evaluated four at a
// time and their lane products combined together: v_permlane32_swap pairs
// lanes l and l + 32 of two slots at once, v_permlane16_swap rows 0 / 1 and
// 2 / 3 of the two pairs (products, so each lane then holds the factors of
// four lanes of one slot: one log2 per lane for four slots), then a row scan
// (row_shr 1, 2, 4, 8) of the logarithms leaves slot t + {0, 2, 1, 3} in
// lane 15 of rows 0..3.  Each slot's value is
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

// the kernel's table: entry idx = (T[idx], T[idx - D]) (0 below D: never read)
__global__ __launch_bounds__(256) void k_rbt_pair(int64_t n, int D, const float *T, float2 *T2) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    T2[t] = make_float2(T[t], t >= D ? T[t - D] : 0.0f);
}

// theta in FP64: b_h + sum_k W_hk x_k of BF16 weights (and an FP32 bias) is
// exact there, so the value after any sequence of moves equals the value
// computed from scratch (carried rows are bit-identical to re-evaluated ones)
template <int CP, int CS, bool PK>
struct RbtState {
    static constexpr int U = 2 * CP + CS;
    double th[U];
    float p[U], q[U];
};

// the lane's registers: byte offsets of its chunks, biases, the padded ones
template <int CP, int CS, bool PK>
struct RbtLane {
    int off[CP + CS];
    uint32_t pad;
};

// the layout tag of carried rows written by the tied kernel: [K 64 slot
// weights (FP32)][U 64 theta (FP64)][final spins, 16-byte rounded]
#define RBT_CARRY_TAG 4096

__device__ __forceinline__ float2 rbt_ld2(const unsigned char *T2, int byte) {
    return *(const float2 *)(T2 + byte);
}
__device__ __forceinline__ float rbt_ld1(const unsigned char *T2, int byte) {
    return *(const float *)(T2 + byte);
}

// p, q of the lane's units; padded units get p = 0, q = 1
template <int CP, int CS, bool PK>
__device__ __forceinline__ void rbt_pq(RbtState<CP, CS, PK> &S, uint32_t pad) {
#pragma unroll
    for (int u = 0; u < 2 * CP + CS; ++u) {
        const float th = (float)S.th[u];
        const float e = __expf(-2.0f * fabsf(th));
        const float inv = __builtin_amdgcn_rcpf(1.0f + e);
        const float big = inv, small = e * inv;
        const bool pd = (pad >> u) & 1u;
        S.p[u] = pd ? 0.0f : (th >= 0.0f ? big : small);
        S.q[u] = pd ? 1.0f : (th >= 0.0f ? small : big);
    }
}

template <int CP, int CS, bool PK>
__device__ void rbt_start(const DevProblem &P, const unsigned char *T2, RbhLds &L,
                          RbtState<CP, CS, PK> &S, const RbtLane<CP, CS, PK> &ln, const uint32_t *sp,
                          int lane) {
    for (int k = lane; k < P.N; k += WAVE) L.x[k] = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1 : -1;
    for (int s = P.n_off + lane; s < P.K * WAVE; s += WAVE) L.w[wpos(s, P.K)] = 0.0f;
    wave_sync();
#pragma unroll
    for (int u = 0; u < 2 * CP + CS; ++u) {
        const int h = P.rbt_hid[(PK && u == 2 * CP ? (lane & 15) : lane) + 64 * u];
        S.th[u] = h >= 0 ? (double)P.rbm_b[h] : 0.0;
    }
    const int wb = 8 * 2 * P.rbt_ncls * P.rbt_F * P.rbt_Q;
    for (int k = 0; k < P.N; ++k) {
        const double xk = (double)L.x[k];
        const int ck = wb + 8 * P.rbt_kcode[k];
#pragma unroll
        for (int c = 0; c < CP; ++c) {
            const float2 w = rbt_ld2(T2, ck + ln.off[c]);
            S.th[2 * c] = __builtin_fma((double)w.x, xk, S.th[2 * c]);
            S.th[2 * c + 1] = __builtin_fma((double)w.y, xk, S.th[2 * c + 1]);
        }
#pragma unroll
        for (int c = 0; c < CS; ++c)
            S.th[2 * CP + c] =
                __builtin_fma((double)rbt_ld1(T2, ck + ln.off[CP + c]), xk, S.th[2 * CP + c]);
    }
    rbt_pq(S, ln.pad);
}

template <int CP, int CS, bool PK>
__device__ __forceinline__ void rbt_move_theta(const DevProblem &P, const unsigned char *T2,
                                               const RbhLds &L, RbtState<CP, CS, PK> &S,
                                               const RbtLane<CP, CS, PK> &ln, int i, int j) {
    const double xi = -2.0 * (double)L.x[i], xj = -2.0 * (double)L.x[j];
    const int wb = 8 * 2 * P.rbt_ncls * P.rbt_F * P.rbt_Q;
    const int ci = wb + 8 * P.rbt_kcode[i], cj = wb + 8 * P.rbt_kcode[j];
#pragma unroll
    for (int c = 0; c < CP; ++c) {
        const float2 wi = rbt_ld2(T2, ci + ln.off[c]), wj = rbt_ld2(T2, cj + ln.off[c]);
        S.th[2 * c] = __builtin_fma(xj, (double)wj.x, __builtin_fma(xi, (double)wi.x, S.th[2 * c]));
        S.th[2 * c + 1] =
            __builtin_fma(xj, (double)wj.y, __builtin_fma(xi, (double)wi.y, S.th[2 * c + 1]));
    }
#pragma unroll
    for (int c = 0; c < CS; ++c) {
        const float wi = rbt_ld1(T2, ci + ln.off[CP + c]), wj = rbt_ld1(T2, cj + ln.off[CP + c]);
        S.th[2 * CP + c] = __builtin_fma(xj, (double)wj,
                                         __builtin_fma(xi, (double)wi, S.th[2 * CP + c]));
    }
    rbt_pq(S, ln.pad);
}

__device__ __forceinline__ float as_f(uint32_t v) { return __uint_as_float(v); }
__device__ __forceinline__ uint32_t as_u(float v) { return __float_as_uint(v); }

// this lane's products of the factors of the four slots at byte bases b0..b3
// (chunk by chunk, the four slots' reads issued together)
template <int CP, int CS, bool PK>
__device__ __forceinline__ void rbt_lane_prod4(const unsigned char *T2, const RbtState<CP, CS, PK> &S,
                                               const RbtLane<CP, CS, PK> &ln, int b0, int b1, int b2,
                                               int b3, float &p0, float &p1, float &p2,
                                               float &p3) {
    p0 = p1 = p2 = p3 = 1.0f;
#pragma unroll
    for (int c = 0; c < CP; ++c) {
        const float2 g0 = rbt_ld2(T2, b0 + ln.off[c]), g1 = rbt_ld2(T2, b1 + ln.off[c]);
        const float2 g2 = rbt_ld2(T2, b2 + ln.off[c]), g3 = rbt_ld2(T2, b3 + ln.off[c]);
        const float pa = S.p[2 * c], qa = S.q[2 * c], pb = S.p[2 * c + 1], qb = S.q[2 * c + 1];
        p0 *= fmaf(pa, g0.x, qa);
        p1 *= fmaf(pa, g1.x, qa);
        p2 *= fmaf(pa, g2.x, qa);
        p3 *= fmaf(pa, g3.x, qa);
        p0 *= fmaf(pb, g0.y, qb);
        p1 *= fmaf(pb, g1.y, qb);
        p2 *= fmaf(pb, g2.y, qb);
        p3 *= fmaf(pb, g3.y, qb);
    }
#pragma unroll
    for (int c = 0; c < (PK ? 0 : CS); ++c) {
        const int o = ln.off[CP + c];
        const float g0 = rbt_ld1(T2, b0 + o), g1 = rbt_ld1(T2, b1 + o);
        const float g2 = rbt_ld1(T2, b2 + o), g3 = rbt_ld1(T2, b3 + o);
        const float pa = S.p[2 * CP + c], qa = S.q[2 * CP + c];
        p0 *= fmaf(pa, g0, qa);
        p1 *= fmaf(pa, g1, qa);
        p2 *= fmaf(pa, g2, qa);
        p3 *= fmaf(pa, g3, qa);
    }
}

// the same factors, one log2 each, summed
template <int CP, int CS, bool PK>
__device__ __forceinline__ float rbt_lane_logs(const unsigned char *T2, const RbtState<CP, CS, PK> &S,
                                               const RbtLane<CP, CS, PK> &ln, int b) {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < CP; ++c) {
        const float2 g = rbt_ld2(T2, b + ln.off[c]);
        acc += __builtin_amdgcn_logf(fmaf(S.p[2 * c], g.x, S.q[2 * c]));
        acc += __builtin_amdgcn_logf(fmaf(S.p[2 * c + 1], g.y, S.q[2 * c + 1]));
    }
#pragma unroll
    for (int c = 0; c < (PK ? 0 : CS); ++c)
        acc += __builtin_amdgcn_logf(
            fmaf(S.p[2 * CP + c], rbt_ld1(T2, b + ln.off[CP + c]), S.q[2 * CP + c]));
    return acc;
}

// every slot weight and the edge states; returns the anti count (the rbh
// kernels' rbh_eval_all with table factors, four slots at a time)
template <int CP, int CS, bool PK>
__device__ int rbt_eval_all(const DevProblem &P, const Tables<float> &tb, const unsigned char *T2,
                            RbhLds &L, const RbtState<CP, CS, PK> &S, const RbtLane<CP, CS, PK> &ln,
                            int lane) {
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
            const float xf = (sc >> 30) & 1 ? xj : xi;   // the canonical end's spin
            L.ent[pos] = make_int2(a, 8 * ((((sc >> 20) & 0x3ff) * 2 + (xf < 0.0f ? 1 : 0)) * FQ +
                                          (sc & 0xfffff)));
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
        float fs = 1.0f;
        if (PK) {   // the packed chunk: unit lane % 16 of this row's slot
            const int lo = (row & 1) ? e23.y : e01.y, hi = (row & 1) ? e23.w : e01.w;
            fs = fmaf(S.p[2 * CP], rbt_ld1(T2, ((row & 2) ? hi : lo) + ln.off[CP]), S.q[2 * CP]);
        }
        float p0, p1, p2, p3;
        rbt_lane_prod4(T2, S, ln, b0, b1, b2, b3, p0, p1, p2, p3);
        // the four lane products paired across lanes first (l with l + 32,
        // then with l + 16: one slot per row), one logarithm per lane, the
        // row scan; a product outside [1e-30, 1e30] in any lane sends the
        // wavefront to one logarithm per factor, summed in the same tree
        const auto q01 = __builtin_amdgcn_permlane32_swap(as_u(p0), as_u(p1), false, false);
        const auto q23 = __builtin_amdgcn_permlane32_swap(as_u(p2), as_u(p3), false, false);
        const float m01 = as_f(q01[0]) * as_f(q01[1]);
        const float m23 = as_f(q23[0]) * as_f(q23[1]);
        const auto qq = __builtin_amdgcn_permlane16_swap(as_u(m01), as_u(m23), false, false);
        const float pr = (as_f(qq[0]) * as_f(qq[1])) * fs;
        float v;
        if (__ballot(!(pr > 1e-30f && pr < 1e30f)) == 0ull) {
            v = __builtin_amdgcn_logf(pr);
        } else {   // rare: factors far from 1
            const float l0 = rbt_lane_logs(T2, S, ln, b0), l1 = rbt_lane_logs(T2, S, ln, b1);
            const float l2 = rbt_lane_logs(T2, S, ln, b2), l3 = rbt_lane_logs(T2, S, ln, b3);
            const auto r01 = __builtin_amdgcn_permlane32_swap(as_u(l0), as_u(l1), false, false);
            const auto r23 = __builtin_amdgcn_permlane32_swap(as_u(l2), as_u(l3), false, false);
            const float s01 = as_f(r01[0]) + as_f(r01[1]);
            const float s23 = as_f(r23[0]) + as_f(r23[1]);
            const auto r = __builtin_amdgcn_permlane16_swap(as_u(s01), as_u(s23), false, false);
            v = as_f(r[0]) + as_f(r[1]);
            if (PK) v += __builtin_amdgcn_logf(fs);
        }
        // rows 0..3: slots t, t + 2, t + 1, t + 3
        v += dpp_get<0x111, 0xf>(v);
        v += dpp_get<0x112, 0xf>(v);
        v += dpp_get<0x114, 0xf>(v);
        v += dpp_get<0x118, 0xf>(v);
        // the slot's log weight in place of its field term; the weights are
        // made after the loop, 64 slots per instruction instead of 4
        if ((lane & 15) == 15) L.g[t + mine] = L.g[t + mine] + 0.69314718055994531f * v;
    }
    wave_sync();
    for (int k = lane; k < n_act; k += WAVE) {
        const int a = L.ent[k].x;
        const float w = __expf(L.g[k]);
        L.w[a & 0xffffff] = (a & RBH_POS_BARE) ? -w : w;
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
template <int CP, int CS, bool PK>
__global__ __launch_bounds__(1024) void k_propagate_rbt(
    DevProblem P, Tables<float> tb, int64_t nw, uint32_t *spins, double *log_w, double *e_loc,
    double beta, double alw, uint32_t key0, uint32_t key1, uint32_t step, uint32_t tag,
    int32_t *evals, const uint32_t *step_base, CarryDev cy) {
    step += step_base ? *step_base : 0u;
    extern __shared__ __align__(16) unsigned char lds[];
    const unsigned char *T2 = lds;
    {
        const uint4 *src = (const uint4 *)P.rbt_tab;
        uint4 *dst = (uint4 *)lds;
        for (int t = threadIdx.x; t < P.rbt_words / 4; t += blockDim.x) dst[t] = src[t];
        __syncthreads();
    }
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    RbhLds L = rbh_carve_at(P, lds, wave, (size_t)P.rbt_words * 4);
    WalkerLds<float> WL;
    WL.w = L.w;
    WL.x = L.x;
    WL.anti = L.anti;
    RbtLane<CP, CS, PK> ln;
#pragma unroll
    for (int c = 0; c < CP + CS; ++c)
        ln.off[c] = 8 * P.rbt_unit[(PK && c == CP ? (lane & 15) : lane) + 64 * c];
    ln.pad = 0u;
#pragma unroll
    for (int u = 0; u < 2 * CP + CS; ++u) {
        if (P.rbt_hid[(PK && u == 2 * CP ? (lane & 15) : lane) + 64 * u] < 0) ln.pad |= 1u << u;
    }
    RbtState<CP, CS, PK> S;
    for (int64_t wk = rbh_next(P.rbm_queue, lane); wk < nw; wk = rbh_next(P.rbm_queue, lane)) {
        uint32_t *sp = spins + wk * P.words;
        // carried rows (gfmc_propagate_carry): the parent's slot weights,
        // theta and anti-parallel count, valid when the rows are this
        // kernel's, of the current parameters, and the parent ended its
        // step with the spins just loaded
        int cur = 0;
        bool carried = false;
        if (cy.on) {
            cur = cy.state[1];
            carried = cy.state[0] != 0 && cy.state[2] == RBT_CARRY_TAG && cy.state[3] == P.rbt_gen;
            if (carried) {
                const uint32_t *cs = (const uint32_t *)((const float *)cy.w[cur] +
                                                        (int64_t)cy.parents[wk] * cy.row +
                                                        cy.s_off);
                bool same = true;
                for (int q = lane; q < P.words; q += WAVE) same = same && cs[q] == sp[q];
                carried = __ballot(!same) == 0ull;
            }
        }
        int n_anti;
        if (carried) {
            const int64_t par = cy.parents[wk];
            const float *row = (const float *)cy.w[cur] + par * cy.row;
            for (int k = lane; k < P.N; k += WAVE) L.x[k] = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1 : -1;
            const uint4 *src = (const uint4 *)row;
            uint4 *dst = (uint4 *)L.w;
            for (int q = lane; q < P.K * 16; q += WAVE) dst[q] = src[q];
            const double *th = (const double *)(row + cy.f_off);
#pragma unroll
            for (int u = 0; u < 2 * CP + CS; ++u) S.th[u] = th[u * WAVE + lane];
            rbt_pq(S, ln.pad);
            n_anti = cy.anti[cur][par];
            wave_sync();
        } else {
            rbt_start(P, T2, L, S, ln, sp, lane);
            n_anti = rbt_eval_all(P, tb, T2, L, S, ln, lane);
        }
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
            if (hop_rate(R)) {
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
                rbt_move_theta(P, T2, L, S, ln, i, j);
                wave_sync();
                if (lane == 0) {
                    L.x[i] = (signed char)-L.x[i];
                    L.x[j] = (signed char)-L.x[j];
                }
                wave_sync();
                n_anti = rbt_eval_all(P, tb, T2, L, S, ln, lane);
                diag = diag_value(P, L.x, L.anti, n_anti, lane);
            }
        }
        lw = lw - alw * beta;
        float *crow = cy.on ? (float *)cy.w[cur ^ 1] + wk * cy.row : nullptr;
        store_spins(P, sp, L.x, lane, cy.on ? (uint32_t *)(crow + cy.s_off) : nullptr);
        if (cy.on) {
            wave_sync();
            const uint4 *src = (const uint4 *)L.w;
            uint4 *dst = (uint4 *)crow;
            for (int q = lane; q < P.K * 16; q += WAVE) dst[q] = src[q];
            double *th = (double *)(crow + cy.f_off);
#pragma unroll
            for (int u = 0; u < 2 * CP + CS; ++u) th[u * WAVE + lane] = S.th[u];
            if (lane == 0) cy.anti[cur ^ 1][wk] = n_anti;
        }
        if (lane == 0) {
            log_w[wk] = lw;
            e_loc[wk] = eL;
            if (evals) evals[wk] += (int32_t)it;
        }
        wave_sync();
    }
}

// after a tied-RBM propagation that stored its rows
__global__ void k_rbt_carry_commit(int32_t *state, int32_t gen) {
    if (threadIdx.x == 0) {
        state[0] = 1;
        state[1] ^= 1;
        state[2] = RBT_CARRY_TAG;
        state[3] = gen;
    }
}
