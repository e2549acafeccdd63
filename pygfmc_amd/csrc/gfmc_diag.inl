// gfmc_diag.inl -- diagonal operators measured by forward walking
// (included at the end of gfmc_hip.hip, one translation unit).
//
// Per production step, after the resampling (driver.py:240-288):
//   k_diag_store  one thread per resampled walker: its parent index and the
//                 value of every operator column go to ring slot t mod R
//                 (R = m_max + 1);
//   k_diag_walk   one thread per walker: follow the stored parents back
//                 through slots t, t-1, ... (the lags m_j ascending, so one
//                 walk serves every window), read the ancestor's value at
//                 slot t - m_j, and reduce over the workgroup in a fixed
//                 order (wavefront butterfly, then the four wave sums in
//                 order) into one partial per (window, column, workgroup);
//   k_diag_finish one workgroup per (window, column): the partials in a
//                 fixed order, divided by n_walkers, into the step's row;
//                 workgroup 0 advances the ring's counter t.
// The counter lives in the ring (device memory), so a captured graph of
// steps continues it on every replay.  Ring layout (bytes): [0, 256) the
// int64 counter, then R * nw int32 parents, then (256-aligned) R * nw *
// n_cols FP64 values, slot-major.

#define DG_T 256
#define DG_MAX_COLS 256
#define DG_MAX_P 64
#define DG_UNIFORM 1   // every edge coefficient of the column is equal
#define DG_LINEAR 2    // the column has a non-zero site term

struct DevDiag {
    int N, words, V, n_p, R;
    const int32_t *kind, *flags, *eoff, *ei, *ej, *m;
    const double *ec, *c0, *a, *b, *c;
};

struct gfmc_diag {
    DevDiag dd;
    int m_max = 0;
    std::vector<void *> allocs;
};

static int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

static int64_t diag_par_offset() { return 256; }

static int64_t diag_val_offset(const gfmc_diag *d, int64_t nw) {
    return 256 + align256((int64_t)d->dd.R * nw * 4);
}

static int64_t diag_row_doubles(const gfmc_diag *d) { return (int64_t)d->dd.V * d->dd.n_p; }

static int64_t diag_blocks(int64_t nw) { return (nw + DG_T - 1) / DG_T; }

__device__ __forceinline__ bool dg_up(const uint32_t *sp, int k) {
    return (sp[k >> 5] >> (k & 31)) & 1u;
}

// value of column v on the packed walker sp (summation in site / edge order)
__device__ double dg_value(const DevDiag &D, const uint32_t *sp, int v) {
    const double *a = D.a + (int64_t)v * D.N;
    if (D.kind[v] == GFMC_DIAG_SF) {
        const double *b = D.b + (int64_t)v * D.N;
        double C = 0.0, S = 0.0;
        for (int q = 0; q < D.words; ++q) {
            const uint32_t w = sp[q];
            const int kmax = min(32, D.N - 32 * q);
            for (int bit = 0; bit < kmax; ++bit) {
                const int k = 32 * q + bit;
                const bool up = (w >> bit) & 1u;
                C += up ? a[k] : -a[k];
                S += up ? b[k] : -b[k];
            }
        }
        return 0.25 * (C * C + S * S) - D.c[v];
    }
    const int e0 = D.eoff[v], e1 = D.eoff[v + 1];
    const int fl = D.flags[v];
    double val;
    if (fl & DG_UNIFORM) {
        int cnt = 0;
        for (int e = e0; e < e1; ++e) cnt += (dg_up(sp, D.ei[e]) == dg_up(sp, D.ej[e])) ? 1 : -1;
        val = D.c0[v] * (double)cnt;
    } else {
        val = 0.0;
        for (int e = e0; e < e1; ++e)
            val += (dg_up(sp, D.ei[e]) == dg_up(sp, D.ej[e])) ? D.ec[e] : -D.ec[e];
    }
    if (fl & DG_LINEAR) {
        double s = 0.0;
        for (int q = 0; q < D.words; ++q) {
            const uint32_t w = sp[q];
            const int kmax = min(32, D.N - 32 * q);
            for (int bit = 0; bit < kmax; ++bit) {
                const int k = 32 * q + bit;
                s += ((w >> bit) & 1u) ? a[k] : -a[k];
            }
        }
        val += s;
    }
    return val + D.c[v];
}

// sum over the DG_T threads of a workgroup, valid in thread 0; fixed order
__device__ __forceinline__ double dg_block_sum(double x, double *red) {
    for (int d = WAVE / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, WAVE);
    const int wv = threadIdx.x / WAVE;
    __syncthreads();
    if ((threadIdx.x & (WAVE - 1)) == 0) red[wv] = x;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < DG_T / WAVE; ++i) s += red[i];
    return s;
}

__device__ __forceinline__ int dg_slot(int64_t t, int64_t back, int R) {
    int64_t r = (t - back) % R;
    return (int)(r < 0 ? r + R : r);
}

__global__ __launch_bounds__(DG_T) void k_diag_ring_init(int64_t nw, int R, int V,
                                                         int64_t *hdr, int32_t *par,
                                                         double *vals) {
    const int64_t i = blockIdx.x * (int64_t)DG_T + threadIdx.x;
    if (i == 0) hdr[0] = 0;
    if (i < (int64_t)R * nw) {
        par[i] = (int32_t)(i % nw);
        for (int v = 0; v < V; ++v) vals[i * V + v] = 0.0;
    }
}

__global__ __launch_bounds__(DG_T) void k_diag_values(DevDiag D, int64_t nw,
                                                      const uint32_t *spins, double *out) {
    const int64_t w = blockIdx.x * (int64_t)DG_T + threadIdx.x;
    if (w >= nw) return;
    const uint32_t *sp = spins + w * D.words;
    for (int v = 0; v < D.V; ++v) out[w * D.V + v] = dg_value(D, sp, v);
}

__global__ __launch_bounds__(DG_T) void k_diag_store(DevDiag D, int64_t nw,
                                                     const uint32_t *spins,
                                                     const int32_t *parents,
                                                     const int64_t *hdr, int32_t *par,
                                                     double *vals) {
    const int64_t w = blockIdx.x * (int64_t)DG_T + threadIdx.x;
    if (w >= nw) return;
    const int slot = dg_slot(hdr[0], 0, D.R);
    par[(int64_t)slot * nw + w] = parents[w];
    const uint32_t *sp = spins + w * D.words;
    double *out = vals + ((int64_t)slot * nw + w) * D.V;
    for (int v = 0; v < D.V; ++v) out[v] = dg_value(D, sp, v);
}

__global__ __launch_bounds__(DG_T) void k_diag_walk(DevDiag D, int64_t nw, const int64_t *hdr,
                                                    const int32_t *par, const double *vals,
                                                    double *part, int64_t nb) {
    __shared__ double red[DG_T / WAVE];
    const int64_t w = blockIdx.x * (int64_t)DG_T + threadIdx.x;
    const bool live = w < nw;
    const int64_t t = hdr[0];
    int64_t anc = live ? w : 0;
    int k = 0;
    for (int j = 0; j < D.n_p; ++j) {
        const int mj = D.m[j];
        for (; k < mj; ++k)
            if (live) anc = par[(int64_t)dg_slot(t, k, D.R) * nw + anc];
        const double *row = vals + ((int64_t)dg_slot(t, mj, D.R) * nw + anc) * D.V;
        for (int v = 0; v < D.V; ++v) {
            const double s = dg_block_sum(live ? row[v] : 0.0, red);
            if (threadIdx.x == 0) part[((int64_t)j * D.V + v) * nb + blockIdx.x] = s;
        }
    }
}

__global__ __launch_bounds__(DG_T) void k_diag_finish(int V, int n_p, int64_t nw, int64_t nb,
                                                      const double *part, double *row,
                                                      int64_t *hdr) {
    __shared__ double red[DG_T / WAVE];
    const int idx = blockIdx.x;   // j * V + v
    const double *p = part + (int64_t)idx * nb;
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += DG_T) s += p[b];
    s = dg_block_sum(s, red);
    if (threadIdx.x == 0) {
        const int j = idx / V, v = idx % V;
        row[(int64_t)v * n_p + j] = s / (double)nw;
        if (idx == 0) hdr[0] += 1;   // no workgroup of this kernel reads the counter
    }
}

static int diag_check(const char *fn, const gfmc_diag *d, const gfmc_problem *p, void *ring,
                      double *dwork) {
    if (!d) return fail(GFMC_EINVAL, std::string(fn) + ": NULL diagonal operators");
    if (!ring || !dwork) return fail(GFMC_EINVAL, std::string(fn) + ": NULL ring or workspace");
    if (p && p->dp.N != d->dd.N)
        return fail(GFMC_EINVAL, std::string(fn) + ": the diagonal operators have " +
                                     std::to_string(d->dd.N) + " sites, the problem " +
                                     std::to_string(p->dp.N));
    return 0;
}

static int diag_record_impl(const gfmc_diag *d, int64_t nw, const uint32_t *spins,
                            const int32_t *parents, void *ring, double *row, double *work,
                            hipStream_t st) {
    const DevDiag &D = d->dd;
    char *base = (char *)ring;
    int64_t *hdr = (int64_t *)base;
    int32_t *par = (int32_t *)(base + diag_par_offset());
    double *vals = (double *)(base + diag_val_offset(d, nw));
    const int64_t nb = diag_blocks(nw);
    hipLaunchKernelGGL(k_diag_store, dim3((unsigned)nb), dim3(DG_T), 0, st, D, nw, spins,
                       parents, (const int64_t *)hdr, par, vals);
    hipLaunchKernelGGL(k_diag_walk, dim3((unsigned)nb), dim3(DG_T), 0, st, D, nw,
                       (const int64_t *)hdr, (const int32_t *)par, (const double *)vals, work,
                       nb);
    hipLaunchKernelGGL(k_diag_finish, dim3((unsigned)(D.V * D.n_p)), dim3(DG_T), 0, st, D.V,
                       D.n_p, nw, nb, (const double *)work, row, hdr);
    HIP_TRY(hipGetLastError());
    return 0;
}

static void free_diag(gfmc_diag *d) {
    for (void *a : d->allocs) (void)hipFree(a);
    delete d;
}

template <typename T>
static int diag_upload(gfmc_diag *d, const T *src, size_t n, const T **dst) {
    void *a = nullptr;
    if (hipMalloc(&a, (n ? n : 1) * sizeof(T)) != hipSuccess)
        return fail(GFMC_ENOMEM, "gfmc_diag_create: hipMalloc failed");
    d->allocs.push_back(a);
    if (n) HIP_TRY(hipMemcpy(a, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = (const T *)a;
    return 0;
}

extern "C" {

int gfmc_diag_create(const gfmc_diag_desc *d, gfmc_diag **out) {
    if (!d || !out) return fail(GFMC_EINVAL, "gfmc_diag_create: NULL argument");
    *out = nullptr;
    const int N = d->n_sites, V = d->n_cols, n_p = d->n_p;
    if (N < 1 || N > 32768) return fail(GFMC_EINVAL, "gfmc_diag_create: n_sites out of range");
    if (V < 1 || V > DG_MAX_COLS)
        return fail(GFMC_EINVAL, "gfmc_diag_create: n_cols must be in 1..256");
    if (n_p < 1 || n_p > DG_MAX_P)
        return fail(GFMC_EINVAL, "gfmc_diag_create: n_p must be in 1..64");
    if (!d->kind || !d->edge_off || !d->a || !d->c || !d->m)
        return fail(GFMC_EINVAL, "gfmc_diag_create: NULL table");
    int m_prev = 0;
    for (int j = 0; j < n_p; ++j) {
        if (d->m[j] < m_prev || d->m[j] > (1 << 20))
            return fail(GFMC_EINVAL, "gfmc_diag_create: m must be non-decreasing in 0..2^20");
        m_prev = d->m[j];
    }
    if (d->edge_off[0] != 0) return fail(GFMC_EINVAL, "gfmc_diag_create: edge_off[0] != 0");
    bool any_sf = false;
    std::vector<int32_t> flags(V, 0);
    std::vector<double> c0(V, 0.0);
    for (int v = 0; v < V; ++v) {
        const int k = d->kind[v];
        if (k != GFMC_DIAG_PAIR && k != GFMC_DIAG_SF)
            return fail(GFMC_EINVAL, "gfmc_diag_create: unknown column kind");
        const int e0 = d->edge_off[v], e1 = d->edge_off[v + 1];
        if (e1 < e0) return fail(GFMC_EINVAL, "gfmc_diag_create: edge_off decreases");
        if (k == GFMC_DIAG_SF) {
            any_sf = true;
            if (e1 != e0) return fail(GFMC_EINVAL, "gfmc_diag_create: SF column with edges");
            continue;
        }
        if (e1 > e0 && (!d->edge_i || !d->edge_j || !d->edge_c))
            return fail(GFMC_EINVAL, "gfmc_diag_create: NULL edge table");
        bool uniform = e1 > e0;
        for (int e = e0; e < e1; ++e) {
            if (d->edge_i[e] < 0 || d->edge_i[e] >= N || d->edge_j[e] < 0 ||
                d->edge_j[e] >= N)
                return fail(GFMC_EINVAL, "gfmc_diag_create: edge site out of range");
            if (d->edge_c[e] != d->edge_c[e0]) uniform = false;
        }
        if (uniform) {
            flags[v] |= DG_UNIFORM;
            c0[v] = d->edge_c[e0];
        }
        for (int s = 0; s < N; ++s)
            if (d->a[(int64_t)v * N + s] != 0.0) flags[v] |= DG_LINEAR;
    }
    if (any_sf && !d->b) return fail(GFMC_EINVAL, "gfmc_diag_create: NULL b table");
    const int n_e = d->edge_off[V];
    gfmc_diag *g = new gfmc_diag;
    DevDiag &D = g->dd;
    D.N = N;
    D.words = (N + 31) / 32;
    D.V = V;
    D.n_p = n_p;
    g->m_max = m_prev;
    D.R = m_prev + 1;
    std::vector<double> zeros;
    if (!any_sf) zeros.assign((size_t)V * N, 0.0);
    int rc = 0;
    if (!rc) rc = diag_upload(g, d->kind, V, &D.kind);
    if (!rc) rc = diag_upload(g, flags.data(), V, &D.flags);
    if (!rc) rc = diag_upload(g, d->edge_off, V + 1, &D.eoff);
    if (!rc) rc = diag_upload(g, d->edge_i, n_e, &D.ei);
    if (!rc) rc = diag_upload(g, d->edge_j, n_e, &D.ej);
    if (!rc) rc = diag_upload(g, d->edge_c, n_e, &D.ec);
    if (!rc) rc = diag_upload(g, c0.data(), V, &D.c0);
    if (!rc) rc = diag_upload(g, d->a, (size_t)V * N, &D.a);
    if (!rc) rc = diag_upload(g, any_sf ? d->b : zeros.data(), (size_t)V * N, &D.b);
    if (!rc) rc = diag_upload(g, d->c, V, &D.c);
    if (!rc) rc = diag_upload(g, d->m, n_p, &D.m);
    if (rc) {
        free_diag(g);
        return rc;
    }
    *out = g;
    return 0;
}

int gfmc_diag_destroy(gfmc_diag *d) {
    if (d) free_diag(d);
    return 0;
}

int64_t gfmc_diag_ring_bytes(const gfmc_diag *d, int64_t nw) {
    if (!d || nw < 0) return -1;
    return diag_val_offset(d, nw) + (int64_t)d->dd.R * nw * d->dd.V * 8;
}

int64_t gfmc_diag_work_doubles(const gfmc_diag *d, int64_t nw) {
    if (!d || nw < 0) return -1;
    return diag_row_doubles(d) * (diag_blocks(nw) > 0 ? diag_blocks(nw) : 1);
}

int gfmc_diag_ring_init(const gfmc_diag *d, int64_t nw, void *ring, void *stream) {
    if (!d || !ring) return fail(GFMC_EINVAL, "gfmc_diag_ring_init: NULL argument");
    if (nw < 0 || nw > 0x7fffffffLL)
        return fail(GFMC_EINVAL, "gfmc_diag_ring_init: n_walkers out of range");
    char *base = (char *)ring;
    const int64_t n = (int64_t)d->dd.R * nw;
    hipLaunchKernelGGL(k_diag_ring_init, dim3(grid1d(n > 0 ? n : 1, DG_T)), dim3(DG_T), 0,
                       (hipStream_t)stream, nw, d->dd.R, d->dd.V, (int64_t *)base,
                       (int32_t *)(base + diag_par_offset()),
                       (double *)(base + diag_val_offset(d, nw)));
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_diag_values(const gfmc_diag *d, int64_t nw, const uint32_t *spins, double *values,
                     void *stream) {
    if (!d) return fail(GFMC_EINVAL, "gfmc_diag_values: NULL diagonal operators");
    if (nw < 0 || nw > 0x7fffffffLL)
        return fail(GFMC_EINVAL, "gfmc_diag_values: n_walkers out of range");
    if (nw == 0) return 0;
    if (!spins || !values) return fail(GFMC_EINVAL, "gfmc_diag_values: NULL buffer");
    hipLaunchKernelGGL(k_diag_values, dim3(grid1d(nw, DG_T)), dim3(DG_T), 0,
                       (hipStream_t)stream, d->dd, nw, spins, values);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_diag_record(const gfmc_diag *d, int64_t nw, const uint32_t *spins,
                     const int32_t *parents, void *ring, double *row, double *work,
                     void *stream) {
    int rc = diag_check("gfmc_diag_record", d, nullptr, ring, work);
    if (rc) return rc;
    if (nw < 1 || nw > 0x7fffffffLL)
        return fail(GFMC_EINVAL, "gfmc_diag_record: n_walkers out of range");
    if (!spins || !parents || !row) return fail(GFMC_EINVAL, "gfmc_diag_record: NULL buffer");
    return diag_record_impl(d, nw, spins, parents, ring, row, work, (hipStream_t)stream);
}

int gfmc_run_steps_diag(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                        double *e_loc, uint32_t *spins_b, double *e_loc_b, int32_t *parents,
                        double *work, int32_t n_steps, double beta, double alw, uint64_t seed,
                        uint32_t step0, uint32_t tag_prop, uint32_t tag_reconf, double *series,
                        int32_t *evals, const gfmc_diag *d, void *ring, double *diag_series,
                        double *diag_work, void *stream) {
    if (!d) return fail(GFMC_EINVAL, "gfmc_run_steps_diag: NULL diagonal operators");
    return run_steps_impl("gfmc_run_steps_diag", p, nw, spins, log_w, e_loc, spins_b, e_loc_b,
                          parents, work, n_steps, beta, alw, seed, step0, tag_prop, tag_reconf,
                          series, evals, d, ring, diag_series, diag_work, stream);
}

int gfmc_graph_capture_steps_diag(const gfmc_problem *p, int64_t nw, uint32_t *spins,
                                  double *log_w, double *e_loc, uint32_t *spins_b,
                                  double *e_loc_b, int32_t *parents, double *work,
                                  int32_t n_steps, double beta, double alw, uint64_t seed,
                                  uint32_t tag_prop, uint32_t tag_reconf, double *series,
                                  int32_t *evals, const gfmc_diag *d, void *ring,
                                  double *diag_series, double *diag_work, uint32_t *counter,
                                  gfmc_graph **out) {
    if (n_steps < 1) return fail(GFMC_EINVAL, "gfmc_graph_capture_steps_diag: n_steps < 1");
    if (!d) return fail(GFMC_EINVAL, "gfmc_graph_capture_steps_diag: NULL diagonal operators");
    return capture_graph("gfmc_graph_capture_steps_diag", counter, [&](void *st) {
        int rc = gfmc_run_steps_diag(p, nw, spins, log_w, e_loc, spins_b, e_loc_b, parents,
                                     work, n_steps, beta, alw, seed, 0u, tag_prop, tag_reconf,
                                     series, evals, d, ring, diag_series, diag_work, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(WAVE), 0, (hipStream_t)st, counter,
                           (uint32_t)n_steps);
        HIP_TRY(hipGetLastError());
        return 0;
    }, out);
}

}  // extern "C"
