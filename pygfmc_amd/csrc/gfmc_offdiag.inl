// Off-diagonal forward walking (the pure estimator of an operator with
// off-diagonal elements): the operator application (apply_operator_one,
// evolution.py:848-889) and the shadow chain of forward_walk_operator
// (evolution.py:892-1101), included at the end of gfmc_hip.hip.
//
// A measurement at main-chain position t applies the operator O to every
// post-resampling walker x: the guide-transformed rates r_k(x) of O (slot 0
// the diagonal element, then the off-diagonal slots, exactly as gfmc_rates
// returns them) give the walker's value v = sum_k r_k and its successor,
// slot k drawn with probability r_k / v by the propagator's blocked
// inverse-CDF sampler from Philox (0, walker, t, tag_op).  A value v <= 0
// leaves the configuration with value 0; an operator without off-diagonal
// slots keeps the configuration and takes its (signed) diagonal element.
// The operator-applied ensemble (log weights 0) then runs m_max ordinary
// branching steps at positions t + 1 .. t + m_max with the production
// tags -- the main chain's own future draws, so from a nearly identical
// ensemble the shadow replays the main chain (correlated sampling).  After
// every shadow step the values follow their walkers through the parent map
// and their mean is recorded with the step's log G; the outputs at the lags
// m_j are fw_j = mean v after m_j steps and norm_j = sum of the first m_j
// shadow log G (one thread, ascending, from 0.0).

#define OD_MAX_P 64

struct OdPick {
    int n_p;
    int m[OD_MAX_P];
};

static size_t od_align(size_t b) { return (b + 255) / 256 * 256; }

struct OdLayout {
    size_t spins_a, spins_b, log_w, e_a, e_b, parents, rwork, v_a, v_b, part, rows, rates, total;
};

static OdLayout od_layout(const gfmc_problem *h, const gfmc_problem *op, int64_t nw, int m_max) {
    OdLayout L;
    const size_t n = (size_t)nw;
    const size_t words = (size_t)h->dp.words;
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o += od_align(bytes);
        return at;
    };
    L.spins_a = take(n * words * sizeof(uint32_t));
    L.spins_b = take(n * words * sizeof(uint32_t));
    L.log_w = take(n * sizeof(double));
    L.e_a = take(n * sizeof(double));
    L.e_b = take(n * sizeof(double));
    L.parents = take(n * sizeof(int32_t));
    L.rwork = take((size_t)gfmc_reconfigure_work_doubles(nw) * sizeof(double));
    L.v_a = take(n * sizeof(double));
    L.v_b = take(n * sizeof(double));
    L.part = take((size_t)((nw + DG_T - 1) / DG_T) * sizeof(double));
    L.rows = take((size_t)(m_max + 1) * 2 * sizeof(double));
    L.rates = take(op ? n * (size_t)(op->dp.n_off + 1) * sizeof(double) : 0);
    L.total = o;
    return L;
}

// One wavefront per walker: the rates row in the propagator's blocked LDS
// layout, the blocked scan, the draw, the successor.  M = 1 + n_off slots.
__global__ __launch_bounds__(WAVE) void k_apply_op(DevProblem Pop, int64_t nw, int M, int K,
                                                   const double *rates, const uint32_t *spins,
                                                   uint32_t k0, uint32_t k1, uint32_t step,
                                                   uint32_t tag, uint32_t *out_spins,
                                                   double *vals, double *log_w, double *e_loc) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x;
    const int64_t wk = blockIdx.x;
    if (wk >= nw) return;
    DevProblem S;   // the sampler's view: M slots, blocked K per lane
    S.n_off = M;
    S.K = K;
    S.n_hx = 0;
    S.n_pad = 0;
    WalkerLds<double> L = carve<double>(S, lds);
    const double *row = rates + wk * (int64_t)M;
    for (int k = 0; k < K; ++k) {
        const int s = lane * K + k;
        L.w[k * WAVE + lane] = s < M ? row[s] : 0.0;
    }
    __syncthreads();
    int idx = 0;
    double val;
    if (M == 1) {
        val = row[0];
    } else {
        double lane_sum;
        const double incl = lane_scan<double>(S, L, lane, &lane_sum);
        const double R = lane_bcast(incl, WAVE - 1);
        if (R <= 0.0) {
            val = 0.0;
        } else {
            const gfmc_u4 r = gfmc_philox(0u, (uint32_t)wk, step, tag, k0, k1);
            double u = gfmc_u53(r.x, r.y) * R;
            const double cap = below(R);
            if (u > cap) u = cap;
            idx = sample_slot<double>(S, L, incl, lane_sum, u, lane);
            val = R;
        }
    }
    int a = -1, b = -1;
    if (idx > 0) {
        const int s = idx - 1;
        if (s < Pop.n_hx) {
            a = s;
        } else {
            a = Pop.edge_i[s - Pop.n_hx];
            b = Pop.edge_j[s - Pop.n_hx];
        }
    }
    const uint32_t *src = spins + wk * Pop.words;
    uint32_t *dst = out_spins + wk * Pop.words;
    for (int q = lane; q < Pop.words; q += WAVE) {
        uint32_t v = src[q];
        if (a >= 0 && (a >> 5) == q) v ^= 1u << (a & 31);
        if (b >= 0 && (b >> 5) == q) v ^= 1u << (b & 31);
        dst[q] = v;
    }
    if (lane == 0) {
        vals[wk] = val;
        if (log_w) log_w[wk] = 0.0;
        if (e_loc) e_loc[wk] = 0.0;
    }
}

// Start of a shadow chain from given values: copy the walkers, zero log
// weights and local energies, copy the values.
__global__ __launch_bounds__(DG_T) void k_od_start(int64_t nw, int words, const uint32_t *spins,
                                                   const double *vin, uint32_t *sa,
                                                   double *log_w, double *e_loc, double *va) {
    const int64_t i = blockIdx.x * (int64_t)DG_T + threadIdx.x;
    if (i < nw * words) sa[i] = spins[i];
    if (i < nw) {
        log_w[i] = 0.0;
        e_loc[i] = 0.0;
        va[i] = vin[i];
    }
}

// v_out[w] = v_in[parents[w]] (parents NULL: v_in[w]) and the workgroup's
// partial sum of the gathered values in a fixed order.
__global__ __launch_bounds__(DG_T) void k_od_carry(int64_t nw, const int32_t *parents,
                                                   const double *v_in, double *v_out,
                                                   double *part) {
    __shared__ double red[DG_T / WAVE];
    const int64_t w = blockIdx.x * (int64_t)DG_T + threadIdx.x;
    double v = 0.0;
    if (w < nw) {
        v = v_in[parents ? parents[w] : w];
        if (v_out) v_out[w] = v;
    }
    const double s = dg_block_sum(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// *mean = (sum of the nb partials, strided per thread then the workgroup's
// fixed-order sum) / nw; *zero = 0 when given.
__global__ __launch_bounds__(DG_T) void k_od_mean(int64_t nw, int64_t nb, const double *part,
                                                  double *mean, double *zero) {
    __shared__ double red[DG_T / WAVE];
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += DG_T) s += part[b];
    s = dg_block_sum(s, red);
    if (threadIdx.x == 0) {
        *mean = s / (double)nw;
        if (zero) *zero = 0.0;
    }
}

// fw_j = rows[m_j].mean, norm_j = 0 + log G_1 + ... + log G_{m_j} (ascending)
__global__ void k_od_pick(OdPick pk, const double *rows, double *fw, double *norm) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = 0.0;
    int s = 0;
    for (int j = 0; j < pk.n_p; ++j) {
        const int m = pk.m[j];
        for (; s < m; ++s) acc = acc + rows[2 * (s + 1)];
        fw[j] = rows[2 * m + 1];
        norm[j] = acc;
    }
}

static int od_check_pair(const char *fn, const gfmc_problem *h, const gfmc_problem *op) {
    if (!h) return fail(GFMC_EINVAL, std::string(fn) + ": NULL Hamiltonian problem");
    if (op && op->dp.N != h->dp.N)
        return fail(GFMC_EINVAL, std::string(fn) + ": the operator has " +
                                     std::to_string(op->dp.N) + " sites, the Hamiltonian " +
                                     std::to_string(h->dp.N));
    return 0;
}

static int od_apply_lds(const gfmc_problem *op, int *K, size_t *bytes) {
    const int M = op->dp.n_off + 1;
    *K = (M + WAVE - 1) / WAVE;
    *bytes = (size_t)(*K) * WAVE * sizeof(double);
    if (*bytes > 65536)
        return fail(GFMC_EUNSUP, "gfmc_apply_operator: at most 8192 operator slots "
                                 "(1 + off-diagonal terms) are supported");
    return 0;
}

static int apply_operator_impl(const gfmc_problem *op, int64_t nw, const uint32_t *spins,
                               uint64_t seed, uint32_t step, uint32_t tag, uint32_t *out_spins,
                               double *vals, double *rates, double *log_w, double *e_loc,
                               hipStream_t st) {
    int K;
    size_t bytes;
    int rc = od_apply_lds(op, &K, &bytes);
    if (rc) return rc;
    rc = gfmc_rates(op, nw, spins, rates, nullptr, nullptr, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_apply_op, dim3((unsigned)nw), dim3(WAVE), bytes, st, op->dp, nw,
                       op->dp.n_off + 1, K, (const double *)rates, spins, (uint32_t)seed,
                       (uint32_t)(seed >> 32), step, tag, out_spins, vals, log_w, e_loc);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" {

int64_t gfmc_apply_operator_work_doubles(const gfmc_problem *op, int64_t n_walkers) {
    if (!op || n_walkers <= 0) return 0;
    return n_walkers * (int64_t)(op->dp.n_off + 1);
}

int gfmc_apply_operator(const gfmc_problem *op, int64_t nw, const uint32_t *spins,
                        uint64_t seed, uint32_t step, uint32_t tag, uint32_t *spins_out,
                        double *vals, double *rates_work, void *stream) {
    if (!op) return fail(GFMC_EINVAL, "gfmc_apply_operator: NULL operator problem");
    int rc = check_walkers("gfmc_apply_operator", op, nw);
    if (rc) return rc;
    if (nw == 0) return 0;
    if (!spins || !spins_out || !vals || !rates_work)
        return fail(GFMC_EINVAL, "gfmc_apply_operator: NULL buffer");
    if (spins_out == spins)
        return fail(GFMC_EINVAL, "gfmc_apply_operator: spins_out must not alias spins");
    return apply_operator_impl(op, nw, spins, seed, step, tag, spins_out, vals, rates_work,
                               nullptr, nullptr, (hipStream_t)stream);
}

int64_t gfmc_fw_operator_work_bytes(const gfmc_problem *h, const gfmc_problem *op,
                                    int64_t n_walkers, int32_t m_max) {
    if (!h || n_walkers <= 0 || m_max < 0) return 0;
    return (int64_t)od_layout(h, op, n_walkers, m_max).total;
}

int gfmc_forward_walk_operator(const gfmc_problem *h, const gfmc_problem *op, int64_t nw,
                               const uint32_t *spins, const double *values, int32_t n_p,
                               const int32_t *m, double beta, double alw, uint64_t seed,
                               uint32_t step, uint32_t tag_op, uint32_t tag_prop,
                               uint32_t tag_reconf, void *work, double *fw, double *norm,
                               void *stream) {
    const char *fn = "gfmc_forward_walk_operator";
    int rc = od_check_pair(fn, h, op);
    if (rc) return rc;
    rc = check_walkers(fn, h, nw);
    if (rc) return rc;
    if (nw == 0) return fail(GFMC_EINVAL, std::string(fn) + ": no walkers");
    if (!spins || !work || !fw || !norm || !m)
        return fail(GFMC_EINVAL, std::string(fn) + ": NULL buffer");
    if (!op && !values)
        return fail(GFMC_EINVAL, std::string(fn) + ": a diagonal measurement (no operator "
                                                   "problem) needs the walkers' values");
    if (n_p < 1 || n_p > OD_MAX_P)
        return fail(GFMC_EINVAL, std::string(fn) + ": n_p must lie in [1, 64]");
    OdPick pk;
    pk.n_p = n_p;
    for (int j = 0; j < n_p; ++j) {
        if (m[j] < 0 || (j > 0 && m[j] < m[j - 1]))
            return fail(GFMC_EINVAL, std::string(fn) + ": m must be non-decreasing and >= 0");
        pk.m[j] = m[j];
    }
    const int m_max = pk.m[n_p - 1];
    const hipStream_t st = (hipStream_t)stream;
    const OdLayout Lo = od_layout(h, op, nw, m_max);
    unsigned char *base = (unsigned char *)work;
    if (((uintptr_t)base) & 255)
        return fail(GFMC_EINVAL, std::string(fn) + ": work must be 256-byte aligned");
    uint32_t *sa = (uint32_t *)(base + Lo.spins_a), *sb = (uint32_t *)(base + Lo.spins_b);
    double *log_w = (double *)(base + Lo.log_w);
    double *ea = (double *)(base + Lo.e_a), *eb = (double *)(base + Lo.e_b);
    int32_t *parents = (int32_t *)(base + Lo.parents);
    double *rwork = (double *)(base + Lo.rwork);
    double *va = (double *)(base + Lo.v_a), *vb = (double *)(base + Lo.v_b);
    double *part = (double *)(base + Lo.part);
    double *rows = (double *)(base + Lo.rows);
    const int words = h->dp.words;
    const int64_t nb = (nw + DG_T - 1) / DG_T;
    if (op) {
        rc = apply_operator_impl(op, nw, spins, seed, step, tag_op, sa, va,
                                 (double *)(base + Lo.rates), log_w, ea, st);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(k_od_start, dim3(grid1d(nw * words, DG_T)), dim3(DG_T), 0, st, nw,
                           words, spins, values, sa, log_w, ea, va);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_od_carry, dim3((unsigned)nb), dim3(DG_T), 0, st, nw,
                       (const int32_t *)nullptr, (const double *)va, (double *)nullptr, part);
    hipLaunchKernelGGL(k_od_mean, dim3(1), dim3(DG_T), 0, st, nw, nb, (const double *)part,
                       rows + 1, rows);
    HIP_TRY(hipGetLastError());
    for (int s = 1; s <= m_max; ++s) {
        const uint32_t pos = step + (uint32_t)s;
        rc = gfmc_propagate(h, nw, sa, log_w, ea, beta, alw, seed, pos, tag_prop, nullptr,
                            stream);
        if (rc) return rc;
        rc = gfmc_reconfigure(nw, log_w, ea, seed, pos, tag_reconf, nullptr, parents,
                              rows + 2 * s, rwork, stream);
        if (rc) return rc;
        rc = gfmc_gather(h->dp.N, nw, sa, ea, parents, sb, eb, log_w, stream);
        if (rc) return rc;
        hipLaunchKernelGGL(k_od_carry, dim3((unsigned)nb), dim3(DG_T), 0, st, nw,
                           (const int32_t *)parents, (const double *)va, vb, part);
        hipLaunchKernelGGL(k_od_mean, dim3(1), dim3(DG_T), 0, st, nw, nb,
                           (const double *)part, rows + 2 * s + 1, (double *)nullptr);
        HIP_TRY(hipGetLastError());
        std::swap(sa, sb);
        std::swap(ea, eb);
        std::swap(va, vb);
    }
    hipLaunchKernelGGL(k_od_pick, dim3(1), dim3(64), 0, st, pk, (const double *)rows, fw, norm);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
