/*
 * gfmc.h -- C-ABI of libgfmc_hip.so, the MI355X (gfx950) GFMC engine.
 *
 * Plain C types only: every buffer is a device pointer owned by the caller
 * (torch tensors' data_ptr() from the Python side), every size is explicit,
 * and every entry point is stream-ordered on the caller's hipStream_t (passed
 * as void*; NULL is the default stream).  Entry points return 0 on success
 * and a negative GFMC_E* code otherwise; gfmc_last_error() then holds a
 * message (per host thread).
 *
 * The reference (PyGFMC, JAX) has no FFI: its seam is Python.  Each entry
 * point below names the reference function whose job it takes over, and the
 * Python package pygfmc_amd binds them with ctypes behind the reference's own
 * API (run / run_gfmc, SpinHalf.create, JastrowSpinHalf, prepare_walkers_sz,
 * Walker).
 *
 * Spin layout in HBM: bit-packed, one walker = n_words = ceil(N/32) uint32
 * words (walker-major, contiguous), bit (k & 31) of word (k >> 5) set <=> spin
 * k is +1 (sigma_k = +1, i.e. S^z = +1/2).
 */
#ifndef GFMC_H
#define GFMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFMC_ABI_VERSION 1

/* status codes */
#define GFMC_OK 0
#define GFMC_EINVAL -1   /* bad argument (shape, size, pointer)          */
#define GFMC_EHIP -2     /* a HIP runtime call failed                     */
#define GFMC_ENOMEM -3   /* the library's own table allocation failed     */
#define GFMC_EUNSUP -4   /* a combination this build does not implement   */

/* arithmetic of the rates/ratios (weights, times and accumulators are FP64) */
#define GFMC_FP64 0
#define GFMC_FP32 1

/* guide kinds */
#define GFMC_GUIDE_NONE 0
#define GFMC_GUIDE_JASTROW 1
#define GFMC_GUIDE_RBM 2
#define GFMC_GUIDE_MLP 3

/* propagators (evolution.py): continuous time (evolve_walker_cont_time,
 * :559-721), single_step (:322-424) and multi_step (:427-556) */
#define GFMC_METHOD_CONT 0
#define GFMC_METHOD_SINGLE 1
#define GFMC_METHOD_MULTI 2

/* Philox stream tags (counter word 3): disjoint streams per purpose */
#define GFMC_TAG_PROPAGATE 0u
#define GFMC_TAG_RECONFIG 1u
#define GFMC_TAG_INIT 2u
#define GFMC_TAG_WARMUP_PROPAGATE 3u
#define GFMC_TAG_WARMUP_RECONFIG 4u
/* operator application of the i-th off-diagonal measurement operator:
 * GFMC_TAG_OPERATOR + i */
#define GFMC_TAG_OPERATOR 16u

/*
 * Problem descriptor: a SpinHalf Hamiltonian (codes/pygfmc/hamiltonian/
 * SpinHalf.py:70-142, stored-coefficient convention of create():144-288) and
 * its guide, lowered to flat host arrays.  The library copies what it needs
 * to the device at gfmc_problem_create; the caller may free these after.
 *
 * Off-diagonal slot layout = SpinHalf.get_rates (SpinHalf.py:316-369) minus
 * the diagonal: slots [0, n_hx) flip site s, slots [n_hx, n_hx + n_jpm) flip
 * both ends of edge s - n_hx.
 */
typedef struct gfmc_problem_desc {
    int32_t n_sites;          /* N                                             */
    int32_t n_edges;          /* Hamiltonian edges                             */
    int32_t n_hx;             /* 0 or N  (SpinHalf.n_hx_static)                */
    int32_t n_jpm;            /* 0 or n_edges (SpinHalf.n_jpm_static)          */
    int32_t has_jz;           /* SpinHalf.has_jz                               */
    int32_t uniform_jz;       /* SpinHalf.uniform_jz                           */
    int32_t has_hz;           /* SpinHalf.has_hz                               */
    int32_t guide;            /* GFMC_GUIDE_*                                  */
    const int32_t *edge_i;    /* [n_edges]                                     */
    const int32_t *edge_j;    /* [n_edges]                                     */
    const double *jz;         /* [n_edges] (uniform: all equal)                */
    const double *jpm;        /* [n_edges]  J_pm   = (J_x+J_y)/4               */
    const double *jpmpm;      /* [n_edges]  J_pmpm = (J_x-J_y)/4               */
    const double *hz;         /* [N] stored -h_z, or NULL when !has_hz         */
    const double *hx_half;    /* [N] stored -h_x/2, or NULL when n_hx == 0     */
    /* Jastrow guide (jastrow_spin_half.py:79-157, build_cache:400-459) */
    int32_t jas_max_deg;
    const int32_t *jas_site_other; /* [N*max_deg]  partner sites (clipped >= 0)  */
    const double *jas_site_J;      /* [N*max_deg]  J = -beta, 0 on padding      */
    const double *jas_h;           /* [N]          one-body field               */
    const int32_t *jas_edge_other; /* [n_edges*2*max_deg] JastrowOpCache.edge_other */
    const double *jas_edge_J;      /* [n_edges*2*max_deg] JastrowOpCache.edge_J    */
    /* RBM guide: log psi = sum_k a_k x_k + sum_h lncosh(b_h + sum_k W_hk x_k) */
    int32_t rbm_hidden;            /* H                                         */
    const float *rbm_W;            /* [H*N] row-major (hidden-major)            */
    const float *rbm_a;            /* [N]                                       */
    const float *rbm_b;            /* [H]                                       */
    /* incremental update lists (built by the host, see pygfmc_amd/problem.py):
     * after move s the slots dep_slots[dep_off[s] : dep_off[s+1]] are
     * re-evaluated and the edges dep_edges[edge_off[s] : edge_off[s+1]]
     * change their parallel/anti-parallel state.  dep_off == NULL selects a
     * full re-evaluation of every slot after every move.                  */
    const int32_t *dep_off;        /* [n_off+1] or NULL                         */
    const int32_t *dep_slots;
    const int32_t *edge_off;       /* [n_off+1] (required when has_jz)          */
    const int32_t *dep_edges;
    /* two-layer MLP guide (no reference counterpart; the reference's network
     * guide is ClusterMLP, wavefunction/cluster_mlp.py:133-439):
     *   log psi = b3 + sum_c w3_c s(b2_c + sum_k W2_ck h_k),
     *   h_k = bf16(s(b1_k + sum_n W1_kn x_n)),
     * s the tabulated tanh (2048 linear pieces on [0, 8), odd, saturated).
     * W1 and b1 are rounded to the layer-1 grid (BF16, nearest even,
     * magnitudes below 2^-9 flushed to 0 and above 64 clamped to +-64), so
     * the layer-1 sums of up to 510 sites are exact 32-bit integers at scale
     * 2^16; W2 to the BF16 grid (magnitudes below 2^-24 flushed to 0); layer
     * 2 runs on MFMA (BF16 inputs, FP32 accumulation).  h1 a multiple of
     * 64, h2 one of 32, 64, 128, 256; n_sites at most 510. */
    int32_t mlp_h1, mlp_h2;
    const float *mlp_W1;           /* [h1*N] row-major                          */
    const float *mlp_b1;           /* [h1]                                      */
    const float *mlp_W2;           /* [h2*h1] row-major                         */
    const float *mlp_b2;           /* [h2]                                      */
    const float *mlp_w3;           /* [h2]                                      */
    const float *mlp_b3;           /* [1]                                       */
} gfmc_problem_desc;

typedef struct gfmc_problem gfmc_problem;   /* opaque, device tables */

typedef struct gfmc_problem_info {
    int32_t n_sites, n_words, n_off, m_total;
    int32_t precision, guide, block_k, full_recompute;
    int64_t lds_bytes;                          /* per walker-wavefront      */
    int32_t lanes;                              /* lanes per walker (64/32/16) */
    int32_t lanes_ok;                           /* bit mask of usable lanes  */
} gfmc_problem_info;

const char *gfmc_version(void);
const char *gfmc_last_error(void);

/* Lower a SpinHalf (+ guide) to device tables.  Replaces the per-call
 * guide_cache / build_cache step (evolution.py:193-202,
 * jastrow_spin_half.py:400-459).  precision: GFMC_FP64 / GFMC_FP32. */
int gfmc_problem_create(const gfmc_problem_desc *desc, int32_t precision,
                        gfmc_problem **out);
int gfmc_problem_destroy(gfmc_problem *p);

/* Lanes per walker of gfmc_propagate / gfmc_run_steps (no reference
 * counterpart; the tuning knob of the propagation kernel): 64 (one walker
 * per wavefront, the default), 32 or 16 (two or four walkers per wavefront,
 * record path only, every move's dependent slots within the group).  The
 * off-diagonal sums then run in blocked order over the group's lanes
 * (K = ceil(n_off / lanes)), so results are bit-identical to the oracle
 * run with the same lane count, and differ from another lane count in the
 * last bits.  gfmc_problem_info.lanes_ok lists the usable values. */
int gfmc_problem_set_lanes(gfmc_problem *p, int32_t lanes);

/* Launch shape of the record-path propagation kernel (no reference
 * counterpart; tuning knobs like the lanes): waves_per_block 1, 2 or 4
 * wavefronts of walkers per workgroup (record paths only; fewer when the
 * walkers' LDS would exceed 64 KiB per workgroup), xcd_remap 1 to map
 * workgroups to walkers so that each of the eight XCDs holds a contiguous
 * range of walkers.  Results do not depend on either (each walker's work is
 * the same).  Default 1, 0. */
int gfmc_problem_set_launch(gfmc_problem *p, int32_t waves_per_block, int32_t xcd_remap);

/* Propagator of gfmc_propagate / gfmc_run_steps / the graphs and batches
 * built from this problem (require_lam, evolution.py:724-741): method
 * GFMC_METHOD_CONT (default; the step argument is beta) or
 * GFMC_METHOD_SINGLE / GFMC_METHOD_MULTI (the step argument is the integer
 * sub-step count n_bra, lam the finite shift, lam >= max H_xx over the
 * reachable configurations).  Discrete methods: every sub-step draws
 * between staying (weight lam - H_xx) and the off-diagonal slots, log_w
 * gains log(lam - H_xx + R) per sub-step, alw * n_bra is subtracted at the
 * end, and e_loc is the local energy of the final configuration (one more
 * evaluation).  The RBM guide propagates in continuous time only. */
int gfmc_problem_set_method(gfmc_problem *p, int32_t method, double lam);
int gfmc_problem_get_info(const gfmc_problem *p, gfmc_problem_info *out);

/* Walker initialisers (walker.py:143-306).  _sz: every walker a Philox-driven
 * Fisher-Yates permutation of n_up up-spins among N sites
 * (prepare_walkers_sz:257-306).  _random: every site an independent fair bit
 * (prepare_walkers_random:207-254).  spins: [n_walkers * n_words]. */
int gfmc_init_walkers_sz(int32_t n_sites, int64_t n_walkers, int32_t n_up,
                         uint64_t seed, uint32_t *spins, void *stream);
int gfmc_init_walkers_random(int32_t n_sites, int64_t n_walkers,
                             uint64_t seed, uint32_t *spins, void *stream);

/* +-1 doubles <-> bit-packed spins (device buffers). */
int gfmc_pack(int32_t n_sites, int64_t n_walkers, const double *config,
              uint32_t *spins, void *stream);
int gfmc_unpack(int32_t n_sites, int64_t n_walkers, const uint32_t *spins,
                double *config, void *stream);

/* Guide-transformed rates of every walker in the SpinHalf layout, FP64 out
 * (transition_rates, evolution.py:205-261 with SpinHalf.get_rates and
 * JastrowSpinHalf.log_diff_from_operator).  rates: [n_walkers * m_total] or
 * NULL; e_loc: [n_walkers] = diag - sum of off-diagonal weights, or NULL;
 * log_diff: [n_walkers * m_total] guide log-ratios (0 in slot 0) or NULL. */
int gfmc_rates(const gfmc_problem *p, int64_t n_walkers, const uint32_t *spins,
               double *rates, double *e_loc, double *log_diff, void *stream);

/* log psi_T of every walker (JastrowSpinHalf.log_psi,
 * jastrow_spin_half.py:299-324; RBM: the network's log-amplitude). */
int gfmc_log_psi(const gfmc_problem *p, int64_t n_walkers, const uint32_t *spins,
                 double *out, void *stream);

/* One continuous-time branching step of every walker, in place
 * (evolve_walker_cont_time, evolution.py:559-721, vmapped at :817-843):
 * hold times ~ Exp(R_x), moves by inverse-CDF over the off-diagonal weights
 * (sample_index, :30-59), log_w -= e_L * tau per hold, then -= alw * beta.
 * The uniforms of hop h of walker w are Philox4x32-10 with counter
 * (h, w, step, tag) and key (seed lo, seed hi): a 53-bit U for the move
 * and a 53-bit psi for the hold (tau = -log1p(-psi) / R_x).
 * evals: [n_walkers], or NULL: each walker's loop iterations = rate
 * evaluations (iter_count of the reference's while_loop) are ADDED to it. */
int gfmc_propagate(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                   double *log_w, double *e_loc, double beta, double alw,
                   uint64_t seed, uint32_t step, uint32_t tag, int32_t *evals,
                   void *stream);

/* Systematic-resampling reconfiguration (reconfiguration.py:9-62): FP64
 * log-sum-exp of log_w, normalised weights, E_n = sum w e_loc, the prefix
 * sum, parents[i] = min(searchsorted_right(cdf, (i+psi)/n), last positive).
 * psi is Philox (0, 0, step, tag) unless psi_opt is non-NULL (a host double).
 * out2 (device, 2 doubles) <- {log G_n - log n, E_n}.  work: caller-owned
 * device scratch of gfmc_reconfigure_work_doubles(n_walkers) doubles (the
 * prefix sum and per-block partials).  Summation order: blocks of 1024
 * walkers, 4 per thread, the wave tree inside a block, block totals in
 * block order. */
int gfmc_reconfigure(int64_t n_walkers, const double *log_w, const double *e_loc,
                     uint64_t seed, uint32_t step, uint32_t tag,
                     const double *psi_opt, int32_t *parents, double *out2,
                     double *work, void *stream);
int64_t gfmc_reconfigure_work_doubles(int64_t n_walkers);

/* gfmc_reconfigure without a workspace argument, and without any device
 * allocation: no counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as gfmc_reconfigure
 * (bit-identical parents and out2); its scratch is a static device array of
 * gfmc_reconfigure_work_doubles(GFMC_RECONFIGURE_STATIC_MAX_WALKERS) doubles
 * (16.9 MB), one per GPU, held until the process ends.  n_walkers above
 * GFMC_RECONFIGURE_STATIC_MAX_WALKERS is rejected with GFMC_EINVAL.  The
 * stream must belong to the current device.  Every call on one GPU shares
 * that GPU's array: calls that overlap in time on one GPU (different streams
 * or host threads without ordering) race. */
#define GFMC_RECONFIGURE_STATIC_MAX_WALKERS 2097152
int gfmc_reconfigure_static(int64_t n_walkers, const double *log_w, const double *e_loc,
                            uint64_t seed, uint32_t step, uint32_t tag,
                            const double *psi_opt, int32_t *parents, double *out2,
                            void *stream);

/* The gather that follows reconfiguration: spins_out[i] = spins_in[parents[i]],
 * e_loc_out[i] = e_loc_in[parents[i]], log_w_out[i] = 0 (NULL skips). */
int gfmc_gather(int32_t n_sites, int64_t n_walkers, const uint32_t *spins_in,
                const double *e_loc_in, const int32_t *parents,
                uint32_t *spins_out, double *e_loc_out, double *log_w_out,
                void *stream);

/* n_steps production (or warm-up) branching steps in one call: per step t,
 * gfmc_propagate at counter step0 + t with tag_prop, gfmc_reconfigure with
 * tag_reconf, gfmc_gather (_make_prod_step_fn, driver.py:194-342, without
 * operators).  The walkers (spins, log_w = 0 after each step, e_loc) are
 * updated in place; spins_b / e_loc_b are same-sized scratch; parents
 * [n_walkers] int32 and work [gfmc_reconfigure_work_doubles(n_walkers)] doubles
 * are scratch.
 * series [2*n_steps] doubles <- {log G_t - log n, E_t} per step.  evals, when
 * non-NULL, [n_walkers] accumulates the rate evaluations of every step (of
 * the walker slot, before resampling).  Launches are enqueued on stream;
 * nothing synchronises. */
int gfmc_run_steps(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                   double *log_w, double *e_loc, uint32_t *spins_b, double *e_loc_b,
                   int32_t *parents, double *work, int32_t n_steps, double beta,
                   double alw, uint64_t seed, uint32_t step0, uint32_t tag_prop,
                   uint32_t tag_reconf, double *series, int32_t *evals, void *stream);

/* Device test hook for the move sampler the propagator uses: for each of
 * n_draws uniforms u_d, the slot drawn from the non-negative weights
 * w[0:n] with the propagator's blocked CDF (sample_index, evolution.py:30-59).
 * weights/uniforms/out are device buffers; precision as in the problem. */
int gfmc_sample_index(int32_t precision, int32_t n, const double *weights,
                      int32_t n_draws, const double *uniforms, int32_t *out,
                      void *stream);

/* Device test hook for the 64-lane inclusive scan every kernel reduction
 * uses (the "wave tree": Hillis-Steele within 16-lane rows, then the row
 * totals; no reference counterpart -- it fixes the summation order that the
 * oracle mirrors).  in/out: 64 device doubles; precision as above. */
int gfmc_wave_scan(int32_t precision, const double *in, double *out, void *stream);

/* Philox4x32-10 block (Salmon et al., SC'11) for the known-answer test:
 * out[4*i..] = philox(ctr[4*i..], key[0..1]); host pointers. */
int gfmc_philox_host(int32_t n, const uint32_t *ctr, const uint32_t *key,
                     uint32_t *out);

/* ---- Batches of independent lattices -------------------------------------
 * Many lattices of uneven size in one launch (no reference counterpart: the
 * reference runs one Hamiltonian per run_gfmc call, driver.py:1148-1547;
 * batching replaces a loop of runs).  Lattice l owns walkers [woff[l],
 * woff[l+1]) in the order given; spins use a common stride of the largest
 * word count (gfmc_batch_get_info), padding words zero.  Lattice l of a
 * batch reproduces bit for bit the single-lattice calls with the seed
 * seed + l; parents are global walker indices; out2 / series rows hold
 * (log G - log n_l, E_l) per lattice. */
typedef struct gfmc_batch gfmc_batch;

/* All problems need one precision, no RBM guide and the same kernel path
 * (gfmc_problem_info.full_recompute aside); n_walkers[l] >= 1.  The batch
 * copies the problems' device descriptors: they must outlive it. */
int gfmc_batch_create(const gfmc_problem *const *problems, int32_t n_lattices,
                      const int64_t *n_walkers, gfmc_batch **out);
int gfmc_batch_destroy(gfmc_batch *b);
/* Lanes per walker for the whole batch (gfmc_problem_set_lanes); 32 and 16
 * also need every lattice's walker count to be a multiple of 64 / lanes.
 * gfmc_batch_lanes_ok returns the bit mask of usable values. */
int gfmc_batch_set_lanes(gfmc_batch *b, int32_t lanes);
int gfmc_batch_lanes_ok(const gfmc_batch *b);
/* total walkers, spin stride (words) and reconfiguration scratch (doubles) */
int gfmc_batch_get_info(const gfmc_batch *b, int64_t *n_total, int32_t *stride,
                        int64_t *work_doubles);

/* gfmc_propagate for every lattice; alw: device [n_lattices] shifts. */
int gfmc_batch_propagate(const gfmc_batch *b, uint32_t *spins, double *log_w,
                         double *e_loc, double beta, const double *alw, uint64_t seed,
                         uint32_t step, uint32_t tag, int32_t *evals, void *stream);

/* gfmc_reconfigure for every lattice's population (segmented, one set of
 * launches); out2: device [n_lattices][2]. */
int gfmc_batch_reconfigure(const gfmc_batch *b, const double *log_w, const double *e_loc,
                           uint64_t seed, uint32_t step, uint32_t tag, int32_t *parents,
                           double *out2, double *work, void *stream);

/* gfmc_run_steps for the batch; series: device [n_steps][n_lattices][2]. */
int gfmc_batch_run_steps(const gfmc_batch *b, uint32_t *spins, double *log_w, double *e_loc,
                         uint32_t *spins_b, double *e_loc_b, int32_t *parents, double *work,
                         int32_t n_steps, double beta, const double *alw, uint64_t seed,
                         uint32_t step0, uint32_t tag_prop, uint32_t tag_reconf,
                         double *series, int32_t *evals, void *stream);

/* ---- Step loops captured into HIP graphs ----------------------------------
 * The production loop of run_gfmc (the chunked lax.scan of _production,
 * driver.py:940-1143) as one HIP graph of n_steps x (propagate, six
 * reconfiguration kernels, gather) launched per chunk.  Every captured
 * kernel takes its Philox position as step + *counter; the graph's last
 * node adds n_steps to the device counter, so consecutive launches continue
 * the chain exactly as eager calls with step0 = 0, n, 2n, ... plus the
 * counter's initial value.  Buffers, beta, shifts, seed and tags are fixed
 * at capture; series receives the n_steps rows of each launch. */
typedef struct gfmc_graph gfmc_graph;

int gfmc_graph_capture_steps(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                             double *log_w, double *e_loc, uint32_t *spins_b,
                             double *e_loc_b, int32_t *parents, double *work,
                             int32_t n_steps, double beta, double alw, uint64_t seed,
                             uint32_t tag_prop, uint32_t tag_reconf, double *series,
                             int32_t *evals, uint32_t *counter, gfmc_graph **out);
int gfmc_graph_capture_batch_steps(const gfmc_batch *b, uint32_t *spins, double *log_w,
                                   double *e_loc, uint32_t *spins_b, double *e_loc_b,
                                   int32_t *parents, double *work, int32_t n_steps,
                                   double beta, const double *alw, uint64_t seed,
                                   uint32_t tag_prop, uint32_t tag_reconf, double *series,
                                   int32_t *evals, uint32_t *counter, gfmc_graph **out);
int gfmc_graph_launch(gfmc_graph *g, void *stream);
int gfmc_graph_destroy(gfmc_graph *g);

/* ---- The p_array energy estimator on the device -------------------------
 * _energy_and_diagonal_estimates (estimators.py:268-305): E_p = sum_i
 * G_p(i) E(i+p) / sum_i G_p(i), log G_p(i) = C[i+p] - C[i], C the cumulative
 * sum of the steps' log G.  Streamed over the series rows a step loop
 * writes: a ring buffer of the last p_max + 1 values of C and, per p, a
 * running maximum with rescaled sums.  state: gfmc_fw_state_doubles device
 * doubles, [0] steps consumed, [2] 1 once a non-finite value was seen; read
 * E_p = num_p / den_p from the last 2 n_p entries ([len - 2 n_p + j] and
 * [len - n_p + j]).  p_list: n_p (1..64) device int32 values in 1..p_max. */
int64_t gfmc_fw_state_doubles(int32_t n_p, int32_t p_max);
int gfmc_fw_init(double *state, int32_t n_p, int32_t p_max, void *stream);
int gfmc_fw_update(const int32_t *p_list, int32_t n_p, int32_t p_max, const double *rows,
                   int32_t n_rows, double *state, void *stream);

/* ---- RBM guide training ------------------------------------------------
 * No counterpart in the reference, which trains guides by NetKet VMC energy
 * minimisation (train_with_netket, generic_netket_training.py:331-473):
 * here the RBM is trained by minimising the variance of the local energy
 * over configurations GFMC generated.  Parameters travel as one flat FP32
 * vector [W (hidden x n_sites, row-major), a (n_sites), b (hidden)] of
 * gfmc_rbm_n_params elements; the problem itself holds W in BF16.  All
 * calls are stream-ordered; a problem whose parameters are being set must
 * not run on another stream at the same time. */

/* H*N + N + H of an RBM problem, -1 otherwise. */
int64_t gfmc_rbm_n_params(const gfmc_problem *p);

/* The problem's parameters from a flat device vector: W rounded to BF16
 * (nearest even), a and b copied, W's column sums recomputed. */
int gfmc_rbm_set_params(gfmc_problem *p, const float *flat, void *stream);

/* Translation-tied RBM on a torus (no counterpart in the reference; the
 * propagation of gfmc_propagate for an RBM guide whose hidden unit
 * h = f n_T + r sees W[h][k] = w_f(k - r)).  The factor of unit h in a pair
 * slot's ratio then comes from a table of (slot class, x_i, f, site i - r)
 * instead of an exponential per factor; moves, weights and energies are
 * those of the untied kernels up to FP32 rounding (the factors are the
 * same; they are multiplied and combined across lanes in another order).
 * Hidden units come in pairs (f, r), (f, r + d) whose table entries lie
 * pair_code apart, so one 8-byte read gives both factors of a pair.
 * Tables (int32, built by pygfmc_amd.torus.tied_tables): lane_off[(n_pair_
 * chunks + n_single_chunks) 64] the table offset f Q + O - code(r) of each
 * lane chunk's (first) unit (padding F Q + O), lane_unit[(2 n_pair_chunks +
 * n_single_chunks) 64] the hidden unit of each lane register (-1: padding;
 * every unit exactly once), slot_code[n_off] code(first site) | class << 20
 * | swapped << 30 (a slot read from its other end), site_code[n_sites],
 * q_site[Q] the site of entry q, q_pair[n_cls Q] the partner site of entry q
 * in class c (-1: none).  Every table read the kernel can make is checked
 * against the table's bounds here.  The problem must have pair moves only
 * with J_pmpm = 0 and at most 512 hidden units, and the chunk shape one the
 * kernel is built for (at most 4 pair and 2 single chunks, 2 pair + single
 * <= 8); otherwise GFMC_EUNSUP and nothing changes.  gfmc_rbm_set_params
 * rebuilds the tables from the rows r = 0: parameters set later must stay
 * tied (the orbit-tied trainers keep them so), or call gfmc_rbm_untie first,
 * which returns the problem to the untied kernels.  theta is held in FP64,
 * where it is exact (BF16 weights), and the kernel carries rows across steps
 * (gfmc_carry_bytes / gfmc_propagate_carry / gfmc_run_steps_carry): slot
 * weights, theta and the final spins of every walker, bit-identical to a
 * re-evaluation; rows of earlier parameters (before gfmc_rbm_set_params) are
 * not used.  A step graph captured before gfmc_rbm_set_params is captured
 * again after it. */
int gfmc_rbm_set_tied(gfmc_problem *p, int32_t n_filters, int32_t n_T, int32_t Q,
                      int32_t n_cls, int32_t pair_code, int32_t n_pair_chunks,
                      int32_t n_single_chunks, const int32_t *lane_off,
                      const int32_t *lane_unit, const int32_t *slot_code,
                      const int32_t *site_code, const int32_t *q_site, const int32_t *q_pair,
                      void *stream);
int gfmc_rbm_untie(gfmc_problem *p);

/* theta[i][h] = b_h + sum_k W_hk x_ik for n packed configurations (BF16
 * MFMA, FP32 accumulation); theta: [n][hidden] FP32 device buffer. */
int gfmc_rbm_theta(const gfmc_problem *p, int64_t n, const uint32_t *spins,
                   float *theta, void *stream);

/* Bytes of device workspace gfmc_rbm_vmin_grad needs for n samples. */
int64_t gfmc_rbm_vmin_work_bytes(const gfmc_problem *p, int64_t n);

/* Gradient of L = (1/n) sum_i (E_i - mean E)^2 over n packed samples,
 * E_i = sum_k rates_k(x_i) with the problem's current parameters:
 * dL/dp = (2/n) sum_i d_i sum_{k>=1} rates_k (O_p(x_i^k) - O_p(x_i)),
 * O = d log psi / dp, d_i = E_i - mean E clamped to +-clip*sqrt(L) when
 * clip > 0.  grad: flat FP32 [n_params]; out2 (device) <- (mean E, L).
 * Deterministic (fixed-order reductions). */
int gfmc_rbm_vmin_grad(const gfmc_problem *p, int64_t n, const uint32_t *spins,
                       float clip, float *grad, double *out2, void *work,
                       void *stream);

/* One Adam step (Kingma & Ba 2015, in torch.optim.Adam's form) on n FP32
 * device parameters with moment buffers m, v; step counts from 1. */
/* Tied parameters: replace every entry of v[0:n] by the mean of its orbit
 * (orbit o = the indices members[offsets[o] .. offsets[o+1]), device int32
 * arrays, summed in that order in FP64; indices outside [0, n) are skipped).
 * Applied to a gradient it trains a translation-invariant guide with the
 * other kernels unchanged.  No counterpart in the reference, whose NetKet
 * trainer differentiates tied parameters directly (the Jastrow guides' tied
 * beta/h, wavefunction/_tied_params.py). */
int gfmc_orbit_mean(int64_t n, float *v, int32_t n_orbits, const int32_t *offsets,
                    const int32_t *members, void *stream);

/* Tied parameters (continued): out[o] = sum of v over orbit o's members;
 * v[p] += alpha * x[orbit of p].  Same orbit tables as gfmc_orbit_mean. */
int gfmc_orbit_sum(int64_t n, const float *v, int32_t n_orbits, const int32_t *offsets,
                   const int32_t *members, float *out, void *stream);
int gfmc_orbit_axpy(int64_t n, float *v, int32_t n_orbits, const int32_t *offsets,
                    const int32_t *members, const float *x, float alpha, void *stream);

/* Natural-gradient (stochastic-reconfiguration) updates of a tied RBM, the
 * SR option of the reference's trainer (generic_netket_training.py:459-461),
 * here for local-energy variance minimisation on GFMC samples.
 * gfmc_rbm_orbit_O: O [n x n_orbits] float, the orbit-summed log-derivatives
 * of each sample (theta: [n x H] float scratch; n <= 65535).
 * gfmc_sr_solve: S = covariance of the rows of O (FP64), then conjugate
 * gradients on (S + eps_rel diag S + eps_abs I) delta = g (at most max_iter
 * iterations, until |r| <= tol |g|), scaled so that sqrt(delta^T S delta),
 * the rms change of log psi over the samples, is at most max_norm (0: no
 * limit); work: gfmc_sr_work_doubles(m) doubles.  warm_start 1 starts the
 * iterations from the unscaled solution a previous call left in the same
 * work buffer (successive natural-gradient steps change S and g little),
 * 0 from delta = 0. */
int gfmc_rbm_orbit_O(const gfmc_problem *p, int64_t n, const uint32_t *spins, int32_t n_orbits,
                     const int32_t *offsets, const int32_t *members, float *O, float *theta,
                     void *stream);
int64_t gfmc_sr_work_doubles(int32_t m);
int gfmc_sr_solve(int64_t n, int32_t m, const float *O, const float *g, double eps_rel,
                  double eps_abs, int32_t max_iter, double tol, double max_norm,
                  int32_t warm_start, float *delta, double *work, void *stream);

int gfmc_adam_step(int64_t n, float *param, const float *grad, float *m, float *v,
                   int64_t step, float lr, float beta1, float beta2, float eps,
                   void *stream);

/* ---- MLP guide training ----------------------------------------------------
 * No counterpart in the reference (see the RBM section above).  Parameters
 * travel as one flat FP32 vector [W1 (h1 x n_sites, row-major), b1 (h1),
 * W2 (h2 x h1, row-major), b2 (h2), w3 (h2), b3] of gfmc_mlp_n_params
 * elements; the problem holds W1, b1 on the layer-1 grid and W2 on the BF16
 * grid (as gfmc_problem_create rounds them).  The gradient
 * call synchronises the caller's stream with the host once (it sizes the
 * candidate rows) and grows a workspace the problem owns. */

/* Totals since the problem's creation of its propagation: out[0] 16-row
 * blocks through the network (each 16 x h1 x h2 multiply-adds on MFMA),
 * out[1] bulk-synchronous hop iterations. */
int gfmc_mlp_counters(const gfmc_problem *p, int64_t *out);
/* A propagation captured into a graph (gfmc_graph_capture_steps*) cannot
 * look at its row count on the host, so each captured step runs hop_cap
 * bulk-synchronous hop iterations (default 512; iterations after the last
 * walker finished are empty launches) and then counts, on the device, a
 * step that still had a walker unfinished.  gfmc_mlp_unfinished: out[0]
 * that count, out[1] the largest number of hop iterations one step has
 * needed, both since the workspace was attached (a caller sizes the cap from
 * out[1] of eager steps and checks out[0] after the replays).  No
 * counterpart in the reference (its step is one XLA scan). */
int gfmc_mlp_set_hop_cap(gfmc_problem *p, int32_t hop_cap);
int gfmc_mlp_unfinished(const gfmc_problem *p, int64_t *out);
/* The MLP guide's propagation workspace is the caller's (no counterpart in
 * the reference, whose guide evaluation is XLA's): per walker the exact
 * layer-1 pre-activations (32-bit integers at scale 2^16), the candidate
 * slots and the hold state, per hop
 * the 16-row block map and one FP32 log psi per row.  gfmc_mlp_work_bytes:
 * the bytes for n_walkers (-1 if not an MLP problem or too large);
 * gfmc_mlp_set_work: attach a 256-byte aligned device buffer of at least
 * that size, good for calls with up to n_walkers walkers on this problem
 * (rates, log psi, propagation, the training gradient) until the next
 * gfmc_mlp_set_work; it zeroes the propagation counters (stream-ordered). */
int64_t gfmc_mlp_work_bytes(const gfmc_problem *p, int64_t n_walkers);
int gfmc_mlp_set_work(gfmc_problem *p, void *work, int64_t bytes, int64_t n_walkers,
                      void *stream);

/* h1*N + h1 + h2*h1 + 2*h2 + 1 of an MLP problem, -1 otherwise. */
int64_t gfmc_mlp_n_params(const gfmc_problem *p);

/* The problem's parameters from a flat device vector (W1, b1 rounded to the
 * layer-1 grid, W2 to the BF16 grid, as gfmc_problem_create rounds them). */
int gfmc_mlp_set_params(gfmc_problem *p, const float *flat, void *stream);

/* Gradient of L = (1/n) sum_i (E_i - mean E)^2 over n packed samples, as
 * gfmc_rbm_vmin_grad: dL/dp = sum_i g_i sum_k r_ik (O_p(x_i^k) - O_p(x_i)),
 * O = d log psi / dp with straight-through BF16 rounding of the layer-1
 * activations and the slopes of the tabulated tanh; g_i = 2 (E_i - mean)/n,
 * clamped to +-clip*sqrt(L) when clip > 0.  Layer-2 products on BF16 MFMA
 * (FP32 operands split hi + lo).  grad: flat FP32 [n_params]; out2 (device)
 * <- (mean E, L).  Deterministic (fixed-order reductions). */
int gfmc_mlp_vmin_grad(const gfmc_problem *p, int64_t n, const uint32_t *spins, float clip,
                       float *grad, double *out2, void *stream);

/* ---- Diagonal operators measured by forward walking ---------------------
 * The diagonal half of the production step (driver.py:240-288,
 * _diag_ring_buffers :810-838): after each step's resampling, every diagonal
 * operator is evaluated on the resampled walkers and stored, with the
 * step's parent indices, in a ring of m_max + 1 slots; for each window p
 * (m = p / 2) every walker's ancestor m steps back is found by following
 * the stored parents, and the walker mean of the ancestors' stored values
 * becomes the step's row.  Slots not yet written hold identity parents and
 * zero values, so a row whose lag reaches before the ring's first step is
 * zero there, as in the reference.
 *
 * An operator is lowered to value columns (a scalar operator is one column,
 * a structure factor one column per wavevector):
 *   GFMC_DIAG_PAIR: sum_e c_e s_i(e) s_j(e) + sum_k a_k s_k + c
 *     (SpinHalf.diagonal_weight, SpinHalf.py:291-314: c_e = J_z/4, a_k =
 *      -h_z/2; when every c_e of a column is equal the value is evaluated as
 *      c_e * (n_parallel - n_antiparallel), the reference's uniform form);
 *   GFMC_DIAG_SF: ((sum_k a_k s_k)^2 + (sum_k b_k s_k)^2) / 4 - c
 *     (SzSzStructureFactor.diagonal_weight, observables.py:162-165: a = cos,
 *      b = sin of k.r, c = N/4);
 * with s = +-1.  Values and means are FP64. */
#define GFMC_DIAG_PAIR 0
#define GFMC_DIAG_SF 1

typedef struct {
    int32_t n_sites;
    int32_t n_cols;            /* 1..256 value columns                       */
    const int32_t *kind;       /* [n_cols] GFMC_DIAG_PAIR / GFMC_DIAG_SF      */
    const int32_t *edge_off;   /* [n_cols + 1] column c owns edges            */
                               /* edge_off[c] .. edge_off[c+1]-1 (PAIR only)  */
    const int32_t *edge_i;     /* [edge_off[n_cols]] sites, 0..n_sites-1      */
    const int32_t *edge_j;
    const double *edge_c;      /* [edge_off[n_cols]]                          */
    const double *a;           /* [n_cols][n_sites]                           */
    const double *b;           /* [n_cols][n_sites] (read for SF columns)     */
    const double *c;           /* [n_cols]                                    */
    int32_t n_p;               /* 1..64 windows                               */
    const int32_t *m;          /* [n_p] lags p / 2, non-decreasing, >= 0      */
} gfmc_diag_desc;

typedef struct gfmc_diag gfmc_diag;

/* Copy the (host) tables to the device.  The handle is read-only after. */
int gfmc_diag_create(const gfmc_diag_desc *d, gfmc_diag **out);
int gfmc_diag_destroy(gfmc_diag *d);

/* Bytes of the caller-owned ring for n_walkers (the step counter, the
 * (m_max + 1) x n_walkers parents, the (m_max + 1) x n_walkers x n_cols
 * values) and doubles of the per-step workspace. */
int64_t gfmc_diag_ring_bytes(const gfmc_diag *d, int64_t n_walkers);
int64_t gfmc_diag_work_doubles(const gfmc_diag *d, int64_t n_walkers);

/* Reset the ring: counter 0, identity parents, zero values. */
int gfmc_diag_ring_init(const gfmc_diag *d, int64_t n_walkers, void *ring, void *stream);

/* values [n_walkers][n_cols] <- every column on every packed walker (the
 * operators' diagonal_weight). */
int gfmc_diag_values(const gfmc_diag *d, int64_t n_walkers, const uint32_t *spins,
                     double *values, void *stream);

/* One step's measurement after resampling: store the values of the
 * resampled walkers spins and their parents, walk the ancestry, write the
 * walker means to row [n_cols][n_p], advance the ring's counter.  work:
 * gfmc_diag_work_doubles doubles.  Deterministic (fixed-order sums). */
int gfmc_diag_record(const gfmc_diag *d, int64_t n_walkers, const uint32_t *spins,
                     const int32_t *parents, void *ring, double *row, double *work,
                     void *stream);

/* gfmc_run_steps with gfmc_diag_record after every step's resampling:
 * diag_series [n_steps][n_cols][n_p]; the ring continues across calls. */
int gfmc_run_steps_diag(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                        double *log_w, double *e_loc, uint32_t *spins_b, double *e_loc_b,
                        int32_t *parents, double *work, int32_t n_steps, double beta,
                        double alw, uint64_t seed, uint32_t step0, uint32_t tag_prop,
                        uint32_t tag_reconf, double *series, int32_t *evals,
                        const gfmc_diag *d, void *ring, double *diag_series,
                        double *diag_work, void *stream);

/* gfmc_graph_capture_steps over gfmc_run_steps_diag (the ring's counter is
 * device state, so replays continue it). */
int gfmc_graph_capture_steps_diag(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                                  double *log_w, double *e_loc, uint32_t *spins_b,
                                  double *e_loc_b, int32_t *parents, double *work,
                                  int32_t n_steps, double beta, double alw, uint64_t seed,
                                  uint32_t tag_prop, uint32_t tag_reconf, double *series,
                                  int32_t *evals, const gfmc_diag *d, void *ring,
                                  double *diag_series, double *diag_work, uint32_t *counter,
                                  gfmc_graph **out);

/*
 * Carried slot weights (the reference's carried rates: initial_carried /
 * gather_carried, engine/evolution.py, driver.py:181-342).  With a carry
 * buffer the step loop ends every propagation by storing each walker's slot
 * weights, and the next propagation starts each walker from its parent's
 * stored weights instead of re-evaluating every slot.  The stored weights
 * are exactly what a full evaluation of the same configuration gives, so
 * the results are bit-identical to runs without a carry.  A row holds the
 * walker's slot weights, its local fields (the FP32 records' local-field
 * form) and the spins it ended the step with; a walker whose spins differ
 * from its parent row's (changed outside these loops) re-evaluates every
 * slot.  Used by the record path (uniform couplings, Jastrow or no guide);
 * 0 bytes otherwise.  The buffer is caller-owned, 256-byte aligned, and
 * starts with its first 256 bytes zero (zeroing them again forgets every
 * row).
 */
int64_t gfmc_carry_bytes(const gfmc_problem *p, int64_t n_walkers);

/* gfmc_propagate with a carry buffer: starts walker w from the stored rows
 * of parents[w] (the last resampling's parents) when the carry holds rows,
 * and stores this propagation's rows. */
int gfmc_propagate_carry(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                         double *log_w, double *e_loc, double beta, double alw, uint64_t seed,
                         uint32_t step, uint32_t tag, int32_t *evals, const int32_t *parents,
                         void *carry, void *stream);

/* gfmc_run_steps_diag (d may be NULL: no diagonal record) with a carry
 * buffer of gfmc_carry_bytes bytes (NULL: none). */
int gfmc_run_steps_carry(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                         double *log_w, double *e_loc, uint32_t *spins_b, double *e_loc_b,
                         int32_t *parents, double *work, int32_t n_steps, double beta,
                         double alw, uint64_t seed, uint32_t step0, uint32_t tag_prop,
                         uint32_t tag_reconf, double *series, int32_t *evals,
                         const gfmc_diag *d, void *ring, double *diag_series,
                         double *diag_work, void *carry, void *stream);

/* gfmc_graph_capture_steps over gfmc_run_steps_carry (the carry's state is
 * device memory, so replays continue it). */
int gfmc_graph_capture_steps_carry(const gfmc_problem *p, int64_t n_walkers, uint32_t *spins,
                                   double *log_w, double *e_loc, uint32_t *spins_b,
                                   double *e_loc_b, int32_t *parents, double *work,
                                   int32_t n_steps, double beta, double alw, uint64_t seed,
                                   uint32_t tag_prop, uint32_t tag_reconf, double *series,
                                   int32_t *evals, const gfmc_diag *d, void *ring,
                                   double *diag_series, double *diag_work, void *carry,
                                   uint32_t *counter, gfmc_graph **out);

/*
 * Off-diagonal forward walking: the pure estimator of an operator O with
 * off-diagonal elements (forward_walk_operator, evolution.py:892-1101).  O
 * is a SpinHalf lowered like a Hamiltonian (gfmc_problem_create) with the
 * run's guide, so gfmc_rates gives its guide-transformed rates r_k(x) =
 * O_{x'x} psi(x') / psi(x) (slot 0 the diagonal element), all >= 0 for a
 * measurement operator.
 */

/* Doubles of the rates workspace of gfmc_apply_operator. */
int64_t gfmc_apply_operator_work_doubles(const gfmc_problem *op, int64_t n_walkers);

/* Apply O to every walker (apply_operator_one, evolution.py:848-889):
 * vals[w] = sum_k r_k(x_w) and spins_out[w] = the successor of slot k drawn
 * with probability r_k / vals[w] (blocked inverse CDF of the propagator over
 * the 1 + n_off slots, u from Philox (0, w, step, tag)); vals <= 0 keeps the
 * configuration with value 0; an operator without off-diagonal slots keeps
 * it with its diagonal element.  spins_out must not alias spins. */
int gfmc_apply_operator(const gfmc_problem *op, int64_t n_walkers, const uint32_t *spins,
                        uint64_t seed, uint32_t step, uint32_t tag, uint32_t *spins_out,
                        double *vals, double *rates_work, void *stream);

/* Bytes of the (256-byte aligned) workspace of gfmc_forward_walk_operator
 * for lags up to m_max (op may be NULL: a diagonal measurement). */
int64_t gfmc_fw_operator_work_bytes(const gfmc_problem *h, const gfmc_problem *op,
                                    int64_t n_walkers, int32_t m_max);

/* One forward-walking measurement of O at main-chain position `step` on the
 * post-resampling walkers `spins` of Hamiltonian h: apply O (tag_op), or,
 * with op == NULL, start from `values` (a diagonal operator's per-walker
 * values, configurations unchanged); run the shadow chain for m[n_p - 1]
 * branching steps of h at positions step + 1, step + 2, ... with the tags
 * tag_prop / tag_reconf (the production tags: the main chain's own future
 * draws), the values following their walkers through each resampling.
 * Outputs (device, [n_p] each, m a HOST array, non-decreasing, n_p <= 64):
 *   fw[j]   = mean of the values after m[j] shadow steps,
 *   norm[j] = sum of the first m[j] shadow steps' log G (0 at m = 0),
 * the pair (val, log_norm) of forward_walk_operator.  Stream-ordered;
 * deterministic (fixed-order sums).  h's method, lanes and precision apply
 * to the shadow steps. */
int gfmc_forward_walk_operator(const gfmc_problem *h, const gfmc_problem *op, int64_t n_walkers,
                               const uint32_t *spins, const double *values, int32_t n_p,
                               const int32_t *m, double beta, double alw, uint64_t seed,
                               uint32_t step, uint32_t tag_op, uint32_t tag_prop,
                               uint32_t tag_reconf, void *work, double *fw, double *norm,
                               void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GFMC_H */
