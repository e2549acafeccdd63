// gfmc_batch.inl -- many independent lattices of uneven size in one launch
// (included at the end of gfmc_hip.hip, one translation unit).
//
// A batch holds the device descriptors of n_lat problems (same precision,
// same kernel path) and the walker segments: lattice l owns walkers
// [woff[l], woff[l + 1]).  Spins of every lattice use a common stride of
// max(words) words.  Propagation is one launch of k_propagate_batch; the
// reconfiguration is the six segmented kernels, each lattice's population
// resampled on its own; the gather uses global parent indices.  Lattice l
// of a batch run reproduces, bit for bit, a single-lattice run with the
// seed seed + l.

struct gfmc_batch {
    int n_lat = 0, precision = 0, fast = 0, fld = 0, stride = 0, lanes = WAVE, method = 0;
    int64_t n_total = 0, lds_bytes = 0, work_doubles = 0;
    int64_t lds_g[3] = {0, 0, 0};   // per wavefront with 64 / 32 / 16 lanes per walker
    int g_ok[3] = {1, 0, 0};
    unsigned rc_blocks = 0;
    DevProblem *d_P = nullptr;
    Tables<double> *d_T64 = nullptr;
    Tables<float> *d_T32 = nullptr;
    int64_t *d_woff = nullptr;
    int32_t *d_boff = nullptr;
    std::vector<int64_t> woff;
    std::vector<void *> allocs;
};

static void free_batch(gfmc_batch *b) {
    for (void *a : b->allocs) (void)hipFree(a);
    delete b;
}

template <typename V>
static int batch_upload(gfmc_batch *b, const V *src, size_t n, V **dst) {
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, n * sizeof(V)));
    b->allocs.push_back(d);
    HIP_TRY(hipMemcpy(d, src, n * sizeof(V), hipMemcpyHostToDevice));
    *dst = (V *)d;
    return 0;
}

static RcSegs batch_segs(const gfmc_batch *b) {
    RcSegs S;
    S.nseg = b->n_lat;
    S.n = b->n_total;
    S.w = b->d_woff;
    S.b = b->d_boff;
    return S;
}

template <typename T, int FAST, int G>
static void launch_batch_g(const gfmc_batch *b, const Tables<T> *Ts, unsigned grid, size_t lds,
                           uint32_t *spins, double *log_w, double *e_loc, double beta,
                           const double *alw, uint64_t seed, uint32_t step, uint32_t tag,
                           int32_t *evals, hipStream_t st) {
    if (b->method == GFMC_METHOD_SINGLE)
        hipLaunchKernelGGL((k_propagate_batch<T, FAST, G, GFMC_METHOD_SINGLE>), dim3(grid),
                           dim3(WAVE), lds, st, b->d_P, Ts, b->n_lat, b->d_woff, b->n_total,
                           b->stride, spins, log_w, e_loc, beta, alw, seed, step, tag, evals,
                           g_step_base);
    else if (b->method == GFMC_METHOD_MULTI)
        hipLaunchKernelGGL((k_propagate_batch<T, FAST, G, GFMC_METHOD_MULTI>), dim3(grid),
                           dim3(WAVE), lds, st, b->d_P, Ts, b->n_lat, b->d_woff, b->n_total,
                           b->stride, spins, log_w, e_loc, beta, alw, seed, step, tag, evals,
                           g_step_base);
    else
        hipLaunchKernelGGL((k_propagate_batch<T, FAST, G, GFMC_METHOD_CONT>), dim3(grid),
                           dim3(WAVE), lds, st, b->d_P, Ts, b->n_lat, b->d_woff, b->n_total,
                           b->stride, spins, log_w, e_loc, beta, alw, seed, step, tag, evals,
                           g_step_base);
}

// the record path (FAST 1: records, 2: their local-field form, every
// lattice having it)
template <typename T, int FAST>
static void launch_batch_rec(const gfmc_batch *b, const Tables<T> *Ts, uint32_t *spins,
                             double *log_w, double *e_loc, double beta, const double *alw,
                             uint64_t seed, uint32_t step, uint32_t tag, int32_t *evals,
                             hipStream_t st) {
    const unsigned n = (unsigned)b->n_total;
    if (b->lanes == 32)
        launch_batch_g<T, FAST, 32>(b, Ts, (n + 1) / 2, (size_t)b->lds_g[1], spins, log_w,
                                    e_loc, beta, alw, seed, step, tag, evals, st);
    else if (b->lanes == 16)
        launch_batch_g<T, FAST, 16>(b, Ts, (n + 3) / 4, (size_t)b->lds_g[2], spins, log_w,
                                    e_loc, beta, alw, seed, step, tag, evals, st);
    else
        launch_batch_g<T, FAST, 64>(b, Ts, n, (size_t)b->lds_bytes, spins, log_w, e_loc, beta,
                                    alw, seed, step, tag, evals, st);
}

template <typename T>
static void launch_propagate_batch(const gfmc_batch *b, const Tables<T> *Ts, uint32_t *spins,
                                   double *log_w, double *e_loc, double beta, const double *alw,
                                   uint64_t seed, uint32_t step, uint32_t tag, int32_t *evals,
                                   hipStream_t st) {
    const unsigned n = (unsigned)b->n_total;
    if (!b->fast) {
        launch_batch_g<T, 0, 64>(b, Ts, n, (size_t)b->lds_bytes, spins, log_w, e_loc, beta,
                                 alw, seed, step, tag, evals, st);
        return;
    }
    if constexpr (std::is_same<T, float>::value) {
        if (b->fld) {
            launch_batch_rec<T, 2>(b, Ts, spins, log_w, e_loc, beta, alw, seed, step, tag, evals,
                                   st);
            return;
        }
    }
    launch_batch_rec<T, 1>(b, Ts, spins, log_w, e_loc, beta, alw, seed, step, tag, evals, st);
}

extern "C" {

int gfmc_batch_create(const gfmc_problem *const *ps, int32_t n_lat, const int64_t *n_walkers,
                      gfmc_batch **out) {
    if (!ps || !n_walkers || !out || n_lat < 1)
        return fail(GFMC_EINVAL, "gfmc_batch_create: bad argument");
    *out = nullptr;
    const int prec = ps[0] ? ps[0]->precision : 0;
    const int fast = ps[0] ? ps[0]->dp.fast : 0;
    int64_t total = 0, lds = 0, work = 0;
    int stride = 1;
    int64_t blocks = 0;
    std::vector<int64_t> woff(n_lat + 1, 0);
    std::vector<int32_t> boff(n_lat + 1, 0);
    for (int l = 0; l < n_lat; ++l) {
        const gfmc_problem *p = ps[l];
        if (!p) return fail(GFMC_EINVAL, "gfmc_batch_create: NULL problem");
        if (p->dp.guide == GFMC_GUIDE_RBM || p->dp.guide == GFMC_GUIDE_MLP)
            return fail(GFMC_EUNSUP, "gfmc_batch_create: network-guided (RBM, MLP) problems do "
                                     "not batch");
        if (p->precision != prec)
            return fail(GFMC_EINVAL, "gfmc_batch_create: every lattice needs one precision");
        if (p->dp.fast != fast)
            return fail(GFMC_EUNSUP, "gfmc_batch_create: every lattice must take the same "
                                     "kernel path (record path or generic)");
        if (p->dp.method != ps[0]->dp.method)
            return fail(GFMC_EINVAL, "gfmc_batch_create: every lattice needs one method");
        if (n_walkers[l] < 1)
            return fail(GFMC_EINVAL, "gfmc_batch_create: every lattice needs >= 1 walker");
        woff[l] = total;
        boff[l] = (int32_t)blocks;
        total += n_walkers[l];
        blocks += (n_walkers[l] + RB - 1) / RB;
        work += n_walkers[l] + 7 * ((n_walkers[l] + RB - 1) / RB) + 1;
        stride = std::max(stride, p->dp.words);
        lds = std::max(lds, p->lds_bytes);
    }
    int64_t lds_g[3] = {lds, 0, 0};
    int g_ok[3] = {1, fast, fast};
    for (int gi = 1; gi < 3; ++gi) {
        const int G = WAVE >> gi, per_wave = WAVE / G;
        for (int l = 0; l < n_lat; ++l) {
            const gfmc_problem *p = ps[l];
            if (!p->dp.g_ok[gi] || n_walkers[l] % per_wave) g_ok[gi] = 0;
            int64_t b1 = gi == 1 ? (prec == GFMC_FP64 ? walker_lds_bytes<double, 32>(p->dp)
                                                      : walker_lds_bytes<float, 32>(p->dp))
                                 : (prec == GFMC_FP64 ? walker_lds_bytes<double, 16>(p->dp)
                                                      : walker_lds_bytes<float, 16>(p->dp));
            lds_g[gi] = std::max(lds_g[gi], per_wave * b1);
        }
        if (lds_g[gi] > 64 * 1024) g_ok[gi] = 0;
    }
    woff[n_lat] = total;
    boff[n_lat] = (int32_t)blocks;
    if (total > 0x7fffffffLL || blocks > 0x7fffffffLL)
        return fail(GFMC_EINVAL, "gfmc_batch_create: more than 2^31 - 1 walkers");
    gfmc_batch *b = new gfmc_batch;
    b->n_lat = n_lat;
    b->precision = prec;
    b->fast = fast;
    b->fld = fast;
    for (int l = 0; l < n_lat; ++l) b->fld = b->fld && ps[l]->dp.fld;
    b->method = ps[0]->dp.method;
    b->stride = stride;
    b->n_total = total;
    b->lds_bytes = lds;
    b->work_doubles = work;
    b->rc_blocks = (unsigned)blocks;
    b->woff = woff;
    for (int gi = 0; gi < 3; ++gi) {
        b->lds_g[gi] = lds_g[gi];
        b->g_ok[gi] = g_ok[gi];
    }
    std::vector<DevProblem> P(n_lat);
    std::vector<Tables<double>> T64(n_lat);
    std::vector<Tables<float>> T32(n_lat);
    for (int l = 0; l < n_lat; ++l) {
        P[l] = ps[l]->dp;
        T64[l] = ps[l]->t64;
        T32[l] = ps[l]->t32;
    }
    int rc = batch_upload(b, P.data(), P.size(), &b->d_P);
    if (!rc) rc = batch_upload(b, T64.data(), T64.size(), &b->d_T64);
    if (!rc) rc = batch_upload(b, T32.data(), T32.size(), &b->d_T32);
    if (!rc) rc = batch_upload(b, woff.data(), woff.size(), &b->d_woff);
    if (!rc) rc = batch_upload(b, boff.data(), boff.size(), &b->d_boff);
    if (rc) {
        free_batch(b);
        return rc;
    }
    *out = b;
    return 0;
}

int gfmc_batch_destroy(gfmc_batch *b) {
    if (b) free_batch(b);
    return 0;
}

int gfmc_batch_get_info(const gfmc_batch *b, int64_t *n_total, int32_t *stride,
                        int64_t *work_doubles) {
    if (!b) return fail(GFMC_EINVAL, "gfmc_batch_get_info: NULL batch");
    if (n_total) *n_total = b->n_total;
    if (stride) *stride = b->stride;
    if (work_doubles) *work_doubles = b->work_doubles;
    return 0;
}

int gfmc_batch_lanes_ok(const gfmc_batch *b) {
    if (!b) return 0;
    int m = 0;
    for (int gi = 0; gi < 3; ++gi)
        if (b->g_ok[gi]) m |= WAVE >> gi;
    return m;
}

int gfmc_batch_set_lanes(gfmc_batch *b, int32_t lanes) {
    if (!b) return fail(GFMC_EINVAL, "gfmc_batch_set_lanes: NULL batch");
    if (lanes != 64 && lanes != 32 && lanes != 16)
        return fail(GFMC_EINVAL, "gfmc_batch_set_lanes: lanes must be 64, 32 or 16");
    if (!b->g_ok[gidx_of(lanes)])
        return fail(GFMC_EUNSUP, "gfmc_batch_set_lanes: some lattice cannot run with this many "
                                 "lanes per walker (record path, moves within the group, "
                                 "walker counts a multiple of 64 / lanes, LDS)");
    b->lanes = lanes;
    return 0;
}

int gfmc_batch_propagate(const gfmc_batch *b, uint32_t *spins, double *log_w, double *e_loc,
                         double beta, const double *alw, uint64_t seed, uint32_t step,
                         uint32_t tag, int32_t *evals, void *stream) {
    if (!b || !spins || !log_w || !e_loc || !alw)
        return fail(GFMC_EINVAL, "gfmc_batch_propagate: NULL argument");
    if (!(beta > 0) || !std::isfinite(beta))
        return fail(GFMC_EINVAL, "gfmc_batch_propagate: beta must be finite and > 0");
    if (b->precision == GFMC_FP64)
        launch_propagate_batch(b, b->d_T64, spins, log_w, e_loc, beta, alw, seed, step, tag,
                               evals, (hipStream_t)stream);
    else
        launch_propagate_batch(b, b->d_T32, spins, log_w, e_loc, beta, alw, seed, step, tag,
                               evals, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_batch_reconfigure(const gfmc_batch *b, const double *log_w, const double *e_loc,
                           uint64_t seed, uint32_t step, uint32_t tag, int32_t *parents,
                           double *out2, double *work, void *stream) {
    if (!b || !log_w || !e_loc || !parents || !out2 || !work)
        return fail(GFMC_EINVAL, "gfmc_batch_reconfigure: NULL argument");
    launch_reconfigure(batch_segs(b), b->n_total, b->rc_blocks, log_w, e_loc, seed, step, tag, 0,
                       0.0, parents, out2, work, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_batch_run_steps(const gfmc_batch *b, uint32_t *spins, double *log_w, double *e_loc,
                         uint32_t *spins_b, double *e_loc_b, int32_t *parents, double *work,
                         int32_t n_steps, double beta, const double *alw, uint64_t seed,
                         uint32_t step0, uint32_t tag_prop, uint32_t tag_reconf, double *series,
                         int32_t *evals, void *stream) {
    if (!b || !spins || !log_w || !e_loc || !spins_b || !e_loc_b || !parents || !work || !alw ||
        n_steps < 0 || (n_steps > 0 && !series))
        return fail(GFMC_EINVAL, "gfmc_batch_run_steps: bad argument");
    const hipStream_t st = (hipStream_t)stream;
    uint32_t *sa = spins, *sb = spins_b;
    double *ea = e_loc, *eb = e_loc_b;
    const int words = b->stride;
    for (int32_t t = 0; t < n_steps; ++t) {
        int rc = gfmc_batch_propagate(b, sa, log_w, ea, beta, alw, seed, step0 + (uint32_t)t,
                                      tag_prop, evals, stream);
        if (rc) return rc;
        launch_reconfigure(batch_segs(b), b->n_total, b->rc_blocks, log_w, ea, seed,
                           step0 + (uint32_t)t, tag_reconf, 0, 0.0, parents,
                           series + (int64_t)2 * b->n_lat * t, work, st);
        hipLaunchKernelGGL(k_gather, dim3(grid1d(b->n_total * words, 256)), dim3(256), 0, st,
                           words, b->n_total, sa, ea, parents, sb, eb, log_w);
        std::swap(sa, sb);
        std::swap(ea, eb);
    }
    if (sa != spins) {
        HIP_TRY(hipMemcpyAsync(spins, sa, (size_t)b->n_total * words * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(e_loc, ea, (size_t)b->n_total * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
