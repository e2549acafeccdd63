// gfmc_graph.inl -- step loops captured once into a HIP graph and replayed
// (included at the end of gfmc_hip.hip, one translation unit).
//
// The Philox position of every step kernel is step + *counter, where
// counter is a device uint32 the caller owns: a captured loop of n steps
// runs steps *counter .. *counter + n - 1 and its last node adds n to the
// counter, so consecutive replays continue the chain exactly as eager
// gfmc_run_steps calls with step0 = 0, n, 2n, ... would.  Every other
// argument (buffers, beta, shifts, seed) is fixed at capture.  The graph
// holds no memset node (clearing memory inside graphs is left to kernels).

struct gfmc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
};

__global__ void k_step_advance(uint32_t *counter, uint32_t n) {
    if (threadIdx.x == 0) counter[0] += n;
}

static void free_graph(gfmc_graph *g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    delete g;
}

// capture enqueue(stream) on a private stream (capture needs a non-null
// stream), then instantiate
template <typename F>
static int capture_graph(const char *fn, uint32_t *counter, F &&enqueue, gfmc_graph **out) {
    if (!out || !counter) return fail(GFMC_EINVAL, std::string(fn) + ": NULL argument");
    *out = nullptr;
    gfmc_graph *g = new gfmc_graph;
    hipError_t e = hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        free_graph(g);
        return fail(GFMC_EHIP, std::string(fn) + ": " + hipGetErrorString(e));
    }
    g_step_base = counter;
    int rc = enqueue((void *)g->cap);
    g_step_base = nullptr;
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(g->cap, &graph);
    g->graph = graph;
    if (rc) {
        free_graph(g);
        return rc;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        free_graph(g);
        return fail(GFMC_EHIP, std::string(fn) + ": " + hipGetErrorString(e));
    }
    *out = g;
    return 0;
}

extern "C" {

int gfmc_graph_capture_steps(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                             double *e_loc, uint32_t *spins_b, double *e_loc_b,
                             int32_t *parents, double *work, int32_t n_steps, double beta,
                             double alw, uint64_t seed, uint32_t tag_prop, uint32_t tag_reconf,
                             double *series, int32_t *evals, uint32_t *counter,
                             gfmc_graph **out) {
    if (n_steps < 1) return fail(GFMC_EINVAL, "gfmc_graph_capture_steps: n_steps < 1");
    return capture_graph("gfmc_graph_capture_steps", counter, [&](void *st) {
        int rc = gfmc_run_steps(p, nw, spins, log_w, e_loc, spins_b, e_loc_b, parents, work,
                                n_steps, beta, alw, seed, 0u, tag_prop, tag_reconf, series,
                                evals, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(WAVE), 0, (hipStream_t)st, counter,
                           (uint32_t)n_steps);
        HIP_TRY(hipGetLastError());
        return 0;
    }, out);
}

int gfmc_graph_capture_steps_carry(const gfmc_problem *p, int64_t nw, uint32_t *spins,
                                   double *log_w, double *e_loc, uint32_t *spins_b,
                                   double *e_loc_b, int32_t *parents, double *work,
                                   int32_t n_steps, double beta, double alw, uint64_t seed,
                                   uint32_t tag_prop, uint32_t tag_reconf, double *series,
                                   int32_t *evals, const gfmc_diag *d, void *ring,
                                   double *diag_series, double *diag_work, void *carry,
                                   uint32_t *counter, gfmc_graph **out) {
    if (n_steps < 1) return fail(GFMC_EINVAL, "gfmc_graph_capture_steps_carry: n_steps < 1");
    return capture_graph("gfmc_graph_capture_steps_carry", counter, [&](void *st) {
        int rc = gfmc_run_steps_carry(p, nw, spins, log_w, e_loc, spins_b, e_loc_b, parents,
                                      work, n_steps, beta, alw, seed, 0u, tag_prop, tag_reconf,
                                      series, evals, d, ring, diag_series, diag_work, carry, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(WAVE), 0, (hipStream_t)st, counter,
                           (uint32_t)n_steps);
        HIP_TRY(hipGetLastError());
        return 0;
    }, out);
}

int gfmc_graph_capture_batch_steps(const gfmc_batch *b, uint32_t *spins, double *log_w,
                                   double *e_loc, uint32_t *spins_b, double *e_loc_b,
                                   int32_t *parents, double *work, int32_t n_steps, double beta,
                                   const double *alw, uint64_t seed, uint32_t tag_prop,
                                   uint32_t tag_reconf, double *series, int32_t *evals,
                                   uint32_t *counter, gfmc_graph **out) {
    if (n_steps < 1) return fail(GFMC_EINVAL, "gfmc_graph_capture_batch_steps: n_steps < 1");
    return capture_graph("gfmc_graph_capture_batch_steps", counter, [&](void *st) {
        int rc = gfmc_batch_run_steps(b, spins, log_w, e_loc, spins_b, e_loc_b, parents, work,
                                      n_steps, beta, alw, seed, 0u, tag_prop, tag_reconf,
                                      series, evals, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(WAVE), 0, (hipStream_t)st, counter,
                           (uint32_t)n_steps);
        HIP_TRY(hipGetLastError());
        return 0;
    }, out);
}

int gfmc_graph_launch(gfmc_graph *g, void *stream) {
    if (!g || !g->exec) return fail(GFMC_EINVAL, "gfmc_graph_launch: NULL graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return 0;
}

int gfmc_graph_destroy(gfmc_graph *g) {
    if (g) free_graph(g);
    return 0;
}

}  // extern "C"
