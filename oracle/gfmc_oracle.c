/*
 * gfmc_oracle.c -- FP64 CPU restatement of PyGFMC's continuous-time GFMC hot
 * path.  TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels of
 * pygfmc_amd and the CPU baseline timed by bench.py.  No product code links,
 * imports or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository).  The one deliberate structural choice -- shared
 * with the kernels so that move choices are bit-identical given the same
 * uniforms -- is the reduction order of the off-diagonal weight sum and its
 * CDF: "blocked wave order" (blocked_scan below), where slot s belongs to
 * lane s / K (K = ceil(n / 64)), each lane sums its block sequentially and
 * the 64 lane sums are combined by the wave-tree inclusive scan; inside the
 * drawn lane's block the CDF is a tree scan of the block's slots
 * (orc_sample_blocked).  The reference's own order is XLA's (cumsum,
 * evolution.py:633); both are the same sum to within a few ulps.
 * Reconfiguration uses the same idea with 1024-walker blocks (rc_block_incl
 * below).  The continuous-time hold's logarithm is a fixed operation
 * sequence (orc_neg_log_unit) shared with the kernels.
 *
 * Pinning: the rates against the reference's golden local energies and a
 * Kronecker-product matrix, the sampler and resampling against the
 * reference's known-answer tests, Philox against the Random123 vectors
 * (tests/test_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/gfmc.h"

#define WAVE 64

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 bijection with the      */
/* published multipliers and Weyl key increments.                           */
/* ------------------------------------------------------------------------ */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0, 1) from two words. */
double orc_u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

static void draw2(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t step, uint32_t tag,
                  double *u1, double *u2) {
    uint32_t ctr[4] = {c0, c1, step, tag};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    orc_philox(ctr, key, r);
    *u1 = orc_u53(r[0], r[1]);
    *u2 = orc_u53(r[2], r[3]);
}

/* ------------------------------------------------------------------------ */
/* Blocked wave order: lane l owns [l*K, (l+1)*K); lane sums sequential from */
/* 0.0; wave-tree inclusive scan (wave_tree) over the 64 lane sums.         */
/* ------------------------------------------------------------------------ */
/* Lanes per walker G of the propagation kernel (64, 32 or 16): the blocked
 * order runs over G lanes, K = ceil(n / G), and the scan is the first stages
 * of the wave tree (orc_set_lanes; the rates and reconfiguration use 64). */
static int g_lanes = WAVE;

void orc_set_lanes(int lanes) { g_lanes = (lanes == 32 || lanes == 16) ? lanes : WAVE; }

static int block_k(int n) { return n > 0 ? (n + g_lanes - 1) / g_lanes : 1; }

/* Inclusive scan of 64 lane values in "wave tree" order: Hillis-Steele
 * inside each 16-lane row (offsets 1, 2, 4, 8), then rows 1 and 3 add the
 * last lane of the row before, then rows 2 and 3 add lane 31 -- the order in
 * which the kernels' DPP scan (row_shr, row_bcast:15, row_bcast:31) adds. */
static void wave_tree_g(double *s, int G) {
    double t[WAVE];
    for (int d = 1; d < 16; d <<= 1) {
        memcpy(t, s, sizeof(double) * (size_t)G);
        for (int l = 0; l < G; ++l)
            if ((l & 15) >= d) s[l] = t[l] + t[l - d];
    }
    if (G >= 32) {
        memcpy(t, s, sizeof(double) * (size_t)G);
        for (int l = 0; l < G; ++l) {
            int row = l >> 4;
            if (row == 1 || row == 3) s[l] = t[l] + t[row * 16 - 1];
        }
    }
    if (G == WAVE) {
        memcpy(t, s, sizeof t);
        for (int l = 32; l < WAVE; ++l) s[l] = t[l] + t[31];
    }
}

static void wave_tree(double *s) { wave_tree_g(s, WAVE); }

void orc_wave_tree(double *s) { wave_tree(s); }

/* lane_incl[64] <- inclusive scan of the lane sums of v[0:n]; returns total. */
double orc_blocked_scan(const double *v, int n, double *lane_incl) {
    int K = block_k(n), G = g_lanes;
    double s[WAVE];
    for (int l = 0; l < G; ++l) {
        double acc = 0.0;
        for (int k = 0; k < K; ++k) {
            int idx = l * K + k;
            if (idx < n) acc += v[idx];
        }
        s[l] = acc;
    }
    wave_tree_g(s, G);
    if (lane_incl) memcpy(lane_incl, s, sizeof(double) * (size_t)G);
    return s[G - 1];
}

/* The last positive slot of w[0:n] at or before slot lim (a zero-weight
 * slot is never returned, sample_index's rule, evolution.py:30-59); 0 when
 * there is none. */
static int last_positive(const double *w, int n, int lim) {
    for (int k = (lim < n ? lim : n - 1); k >= 0; --k)
        if (w[k] > 0) return k;
    return 0;
}

/* Inverse-CDF draw of a slot of w[0:n] (all >= 0, total > 0) at u in
 * [0, total): j = the first lane whose inclusive sum exceeds u; inside lane
 * j's block of K slots the CDF is base + the group's tree scan (wave_tree_g
 * over G lanes) of G consecutive slots at a time, base starting at the
 * previous lane's inclusive sum and advancing by each chunk's total; the
 * first slot whose CDF exceeds u is drawn.  A zero-weight slot that rounding
 * selects gives way to the last positive slot before it; where rounding
 * leaves u past lane j's last slot, the last positive slot of lane j's block
 * or before it is taken (sample_index, evolution.py:30-59). */
int orc_sample_blocked(const double *w, int n, const double *lane_incl, double u) {
    int K = block_k(n), G = g_lanes;
    int j = 0;
    while (j < G - 1 && !(lane_incl[j] > u)) ++j;
    double base = (j == 0) ? 0.0 : lane_incl[j - 1];
    double s[WAVE];
    for (int c0 = 0; c0 < K; c0 += G) {
        int ok[WAVE];
        for (int l = 0; l < G; ++l) {
            int k = c0 + l, slot = j * K + k;
            ok[l] = k < K && slot < n;
            s[l] = ok[l] ? w[slot] : 0.0;
        }
        wave_tree_g(s, G);
        for (int l = 0; l < G; ++l) {
            if (ok[l] && base + s[l] > u) {
                int slot = j * K + c0 + l;
                return w[slot] > 0 ? slot : last_positive(w, n, slot - 1);
            }
        }
        base = base + s[G - 1];
    }
    int end = j * K + K - 1;
    return last_positive(w, n, end);
}

/* -log(q) for q in [2^-1022, 1] (the continuous-time hold, q = 1 - psi
 * exactly): fdlibm's e_log.c reduction and polynomial (Sun Microsystems,
 * 1993), q = 2^k (1 + f), 1 + f in [sqrt(2)/2, sqrt(2)), s = f / (2 + f),
 * log q = k ln2_hi - ((f^2/2 - (s (f^2/2 + R(s^2)) + k ln2_lo)) - f), the
 * operations of the kernels in the same order. */
double orc_neg_log_unit(double q) {
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double L1 = 6.666666666666735130e-01, L2 = 3.999999999940941908e-01;
    const double L3 = 2.857142874366239149e-01, L4 = 2.222219843214978396e-01;
    const double L5 = 1.818357216161805012e-01, L6 = 1.531383769920937332e-01;
    const double L7 = 1.479819860511658591e-01;
    uint64_t b;
    memcpy(&b, &q, sizeof b);
    uint32_t hi = (uint32_t)(b >> 32) & 0x000fffffu;
    uint32_t up = (hi + 0x95f64u) & 0x100000u;
    int k = (int)((b >> 52) & 0x7ffu) - 1023 + (int)(up >> 20);
    uint64_t mb = ((uint64_t)(hi | (up ^ 0x3ff00000u)) << 32) | (b & 0xffffffffull);
    double m;
    memcpy(&m, &mb, sizeof m);
    double f = m - 1.0;
    double sr = f / (2.0 + f);
    double z = sr * sr, ww = z * z;
    double t1 = ww * (L2 + ww * (L4 + ww * L6));
    double t2 = z * (L1 + ww * (L3 + ww * (L5 + ww * L7)));
    double R = t2 + t1;
    double hfsq = 0.5 * f * f;
    double dk = (double)k;
    return -(dk * ln2_hi - ((hfsq - (sr * (hfsq + R) + dk * ln2_lo)) - f));
}

/* Host entry used by the sampler tests: draws for several uniforms in [0,1)
 * exactly as the propagator does (u = min(U * R, nextafter(R, 0))). */
void orc_sample_many(const double *w, int n, int n_draws, const double *uniforms, int32_t *out) {
    double incl[WAVE];
    double R = orc_blocked_scan(w, n, incl);
    for (int d = 0; d < n_draws; ++d) {
        double u = uniforms[d] * R;
        double cap = nextafter(R, 0.0);
        if (u > cap) u = cap;
        out[d] = orc_sample_blocked(w, n, incl, u);
    }
}

/* ------------------------------------------------------------------------ */
/* SpinHalf rates with the Jastrow guide.                                     */
/* ------------------------------------------------------------------------ */

static int n_off_of(const gfmc_problem_desc *d) { return d->n_hx + d->n_jpm; }

/* Jastrow mean field of a site flip: sum_b J_kb x_b over the site's partner
 * table in slot order (_site_log_diff, jastrow_spin_half.py:290-297). */
static double site_mf(const gfmc_problem_desc *d, const int8_t *x, int k) {
    double acc = 0.0;
    int D = d->jas_max_deg;
    for (int q = 0; q < D; ++q)
        acc += d->jas_site_J[k * D + q] * (double)x[d->jas_site_other[k * D + q]];
    return acc;
}

/* Per-edge external mean field of end `side` (0: i, 1: j) of edge e, the
 * other end excluded (_edge_log_diff_from_cache, jastrow_spin_half.py:462-472,
 * cache from build_cache:400-459). */
static double edge_mf(const gfmc_problem_desc *d, const int8_t *x, int e, int side) {
    double acc = 0.0;
    int D = d->jas_max_deg;
    const int32_t *oth = d->jas_edge_other + ((size_t)e * 2 + side) * D;
    const double *J = d->jas_edge_J + ((size_t)e * 2 + side) * D;
    for (int q = 0; q < D; ++q) acc += J[q] * (double)x[oth[q]];
    return acc;
}

/* ------------------------------------------------------------------------ */
/* RBM guide (no reference counterpart; the reference's network guide is     */
/* ClusterMLP, wavefunction/cluster_mlp.py:133-439): log psi = sum_k a_k x_k  */
/* + sum_h lncosh(theta_h), theta_h = b_h + sum_k W_hk x_k, evaluated in FP64 */
/* from the (BF16-valued) float parameters.                                   */
/* ------------------------------------------------------------------------ */
static double lncosh(double y) {
    double a = fabs(y);
    return a + log1p(exp(-2.0 * a)) - log(2.0);
}

/* theta[h] = b_h + sum_k W_hk x_k, k ascending */
void orc_rbm_theta(const gfmc_problem_desc *d, const int8_t *x, double *theta) {
    int H = d->rbm_hidden, N = d->n_sites;
    for (int h = 0; h < H; ++h) {
        double acc = (double)d->rbm_b[h];
        for (int k = 0; k < N; ++k) acc += (double)d->rbm_W[(size_t)h * N + k] * (double)x[k];
        theta[h] = acc;
    }
}

/* log psi(x') - log psi(x) for flipping site i (and j >= 0) */
static double rbm_log_ratio(const gfmc_problem_desc *d, const int8_t *x, const double *theta,
                            int i, int j) {
    int H = d->rbm_hidden, N = d->n_sites;
    double xi = (double)x[i], xj = j >= 0 ? (double)x[j] : 0.0;
    double lr = -2.0 * (double)d->rbm_a[i] * xi;
    if (j >= 0) lr += -2.0 * (double)d->rbm_a[j] * xj;
    double acc = 0.0;
    for (int h = 0; h < H; ++h) {
        double delta = -2.0 * (double)d->rbm_W[(size_t)h * N + i] * xi;
        if (j >= 0) delta += -2.0 * (double)d->rbm_W[(size_t)h * N + j] * xj;
        acc += lncosh(theta[h] + delta) - lncosh(theta[h]);
    }
    return lr + acc;
}

/* ------------------------------------------------------------------------ */
/* MLP guide (no reference counterpart; nearest ClusterMLP,                  */
/* wavefunction/cluster_mlp.py:133-439):                                      */
/*   log psi = b3 + sum_c w3_c s(b2_c + sum_k W2_ck h_k),                     */
/*   h_k = bf16(s(b1_k + sum_n W1_kn x_n)),                                   */
/* in the library's arithmetic: W1 and b1 on the layer-1 grid (BF16 nearest  */
/* even, magnitudes below 2^-9 flushed, above 64 clamped), W2 on the BF16    */
/* grid (below 2^-24 flushed), layer-1 sums in FP64 (exact on that grid), s the */
/* tabulated tanh in FP32 with an explicit fma, h rounded to BF16; the       */
/* layer-2 sums in FP64 rounded once to FP32 (the device accumulates them in */
/* FP32 on MFMA: the one step not mirrored bit for bit), then the output     */
/* sum in the device's order: per column c mod 16 over c / 16, then a        */
/* butterfly over the 16 columns.                                             */
/* ------------------------------------------------------------------------ */
#define MLP_NT 2048
static float mlp_T[MLP_NT], mlp_D[MLP_NT];
static int mlp_tab_ready = 0;

static void mlp_table(void) {
    if (mlp_tab_ready) return;
    for (int j = 0; j < MLP_NT; ++j) {
        float t0 = (float)tanh((double)j / 256.0), t1 = (float)tanh((double)(j + 1) / 256.0);
        mlp_T[j] = t0;
        mlp_D[j] = t1 - t0;
    }
    mlp_tab_ready = 1;
}

void orc_mlp_table(float *T, float *D) {
    mlp_table();
    memcpy(T, mlp_T, sizeof mlp_T);
    memcpy(D, mlp_D, sizeof mlp_D);
}

/* s(z) = sign(z) (T_j + f D_j), t = min(256 |z|, 2048 - 2^-13), j = floor(t),
 * f = t - j */
float orc_mlp_tanh(float z) {
    mlp_table();
    float t = fminf(fabsf(z) * 256.0f, 2047.9998779296875f);
    int j = (int)t;
    float f = t - floorf(t);
    return copysignf(fmaf(f, mlp_D[j], mlp_T[j]), z);
}

static float bf16_rne(float v) {
    uint32_t b;
    memcpy(&b, &v, 4);
    b = ((b + 0x7fffu + ((b >> 16) & 1u)) >> 16) << 16;
    float r;
    memcpy(&r, &b, 4);
    return r;
}

float orc_bf16_flush(float v) {
    float r = bf16_rne(v);
    return fabsf(r) < 5.9604644775390625e-08f ? 0.0f : r;
}

/* the layer-1 grid of W1 and b1: BF16, |v| < 2^-9 -> 0, |v| > 64 -> +-64 */
float orc_mlp_l1_grid(float v) {
    float r = bf16_rne(v);
    if (fabsf(r) < 0.001953125f) return 0.0f;
    return fabsf(r) > 64.0f ? copysignf(64.0f, r) : r;
}

/* The grid weights of the MLP (W1, b1 after orc_mlp_l1_grid, W2 after
 * orc_bf16_flush), made
 * once per configuration evaluation so the rows of its candidates do not
 * re-round them; the arithmetic is the same as rounding on the fly. */
typedef struct {
    float *W1, *b1, *W2;
} mlp_grid_t;

static size_t mlp_grid_floats(const gfmc_problem_desc *d) {
    return (size_t)d->mlp_h1 * (size_t)d->n_sites + (size_t)d->mlp_h1 +
           (size_t)d->mlp_h2 * (size_t)d->mlp_h1;
}

static mlp_grid_t mlp_grid_fill(const gfmc_problem_desc *d, float *buf) {
    mlp_grid_t g;
    size_t n1 = (size_t)d->mlp_h1 * (size_t)d->n_sites, n2 = (size_t)d->mlp_h2 * d->mlp_h1;
    g.W1 = buf;
    g.b1 = buf + n1;
    g.W2 = buf + n1 + d->mlp_h1;
    for (size_t q = 0; q < n1; ++q) g.W1[q] = orc_mlp_l1_grid(d->mlp_W1[q]);
    for (int k = 0; k < d->mlp_h1; ++k) g.b1[k] = orc_mlp_l1_grid(d->mlp_b1[k]);
    for (size_t q = 0; q < n2; ++q) g.W2[q] = orc_bf16_flush(d->mlp_W2[q]);
    return g;
}

/* layer-1 activations h (BF16 values) of x */
static void mlp_hidden_g(const gfmc_problem_desc *d, const mlp_grid_t *g, const int8_t *x,
                         float *h) {
    int N = d->n_sites, H1 = d->mlp_h1, k = 0;
    /* four hidden units at a time: four independent sums, each in site order */
    for (; k + 4 <= H1; k += 4) {
        const float *w0 = g->W1 + (size_t)k * N, *w1 = w0 + N, *w2 = w1 + N, *w3 = w2 + N;
        double z0 = (double)g->b1[k], z1 = (double)g->b1[k + 1];
        double z2 = (double)g->b1[k + 2], z3 = (double)g->b1[k + 3];
        for (int n = 0; n < N; ++n) {
            const double xn = (double)x[n];
            z0 += xn * (double)w0[n];
            z1 += xn * (double)w1[n];
            z2 += xn * (double)w2[n];
            z3 += xn * (double)w3[n];
        }
        h[k] = bf16_rne(orc_mlp_tanh((float)z0));
        h[k + 1] = bf16_rne(orc_mlp_tanh((float)z1));
        h[k + 2] = bf16_rne(orc_mlp_tanh((float)z2));
        h[k + 3] = bf16_rne(orc_mlp_tanh((float)z3));
    }
    for (; k < H1; ++k) {
        double z = (double)g->b1[k];
        for (int n = 0; n < N; ++n) z += (double)x[n] * (double)g->W1[(size_t)k * N + n];
        h[k] = bf16_rne(orc_mlp_tanh((float)z));
    }
}

static float mlp_log_psi_g(const gfmc_problem_desc *d, const mlp_grid_t *g, const int8_t *x) {
    int H1 = d->mlp_h1, H2 = d->mlp_h2, NC = H2 / 16;
    float *h = (float *)malloc(sizeof(float) * (size_t)H1);
    mlp_hidden_g(d, g, x, h);
    float s[16];
    for (int l = 0; l < 16; ++l) s[l] = 0.0f;
    for (int t = 0; t < NC; ++t)
        for (int l = 0; l < 16; l += 4) {
            /* four columns at a time: four independent sums, each in k order */
            const int c = t * 16 + l;
            const float *w0 = g->W2 + (size_t)c * H1, *w1 = w0 + H1, *w2 = w1 + H1,
                        *w3 = w2 + H1;
            double z[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < H1; ++k) {
                const double hk = (double)h[k];
                z[0] += (double)w0[k] * hk;
                z[1] += (double)w1[k] * hk;
                z[2] += (double)w2[k] * hk;
                z[3] += (double)w3[k] * hk;
            }
            for (int q = 0; q < 4; ++q) {
                float v = d->mlp_w3[c + q] * orc_mlp_tanh((float)z[q] + d->mlp_b2[c + q]);
                s[l + q] = s[l + q] + v;
            }
        }
    for (int dd = 8; dd >= 1; dd >>= 1) {
        float o[16];
        for (int l = 0; l < 16; ++l) o[l] = s[l] + s[l ^ dd];
        memcpy(s, o, sizeof o);
    }
    free(h);
    return s[0] + d->mlp_b3[0];
}

void orc_mlp_hidden(const gfmc_problem_desc *d, const int8_t *x, float *h) {
    float *buf = (float *)malloc(sizeof(float) * mlp_grid_floats(d));
    mlp_grid_t g = mlp_grid_fill(d, buf);
    mlp_hidden_g(d, &g, x, h);
    free(buf);
}

float orc_mlp_log_psi(const gfmc_problem_desc *d, const int8_t *x) {
    float *buf = (float *)malloc(sizeof(float) * mlp_grid_floats(d));
    mlp_grid_t g = mlp_grid_fill(d, buf);
    float r = mlp_log_psi_g(d, &g, x);
    free(buf);
    return r;
}

/* theta of the MLP guide: theta[0] = log psi(x), followed by the grid
 * weights (floats) its candidate rows use */
static mlp_grid_t mlp_grid_of_theta(const gfmc_problem_desc *d, const double *theta) {
    float *buf = (float *)(void *)(theta + 1);
    mlp_grid_t g;
    size_t n1 = (size_t)d->mlp_h1 * (size_t)d->n_sites;
    g.W1 = buf;
    g.b1 = buf + n1;
    g.W2 = buf + n1 + d->mlp_h1;
    return g;
}

static double mlp_log_ratio(const gfmc_problem_desc *d, const int8_t *x, const double *lp0,
                            int i, int j) {
    int N = d->n_sites;
    int8_t *y = (int8_t *)malloc((size_t)N);
    memcpy(y, x, (size_t)N);
    y[i] = (int8_t)-y[i];
    if (j >= 0) y[j] = (int8_t)-y[j];
    const mlp_grid_t g = mlp_grid_of_theta(d, lp0);
    float out = mlp_log_psi_g(d, &g, y);
    free(y);
    return (double)(out - (float)lp0[0]);
}

/* Guide log-ratio of off-diagonal slot s (log_diff_from_operator,
 * jastrow_spin_half.py:369-398); theta: the RBM's theta of x (RBM), or
 * log psi(x) as computed for the MLP guide. */
static double slot_log_diff(const gfmc_problem_desc *d, const int8_t *x, int s,
                            const double *theta) {
    if (d->guide == GFMC_GUIDE_MLP) {
        if (s < d->n_hx) return mlp_log_ratio(d, x, theta, s, -1);
        int e = s - d->n_hx;
        return mlp_log_ratio(d, x, theta, d->edge_i[e], d->edge_j[e]);
    }
    if (d->guide == GFMC_GUIDE_RBM) {
        if (s < d->n_hx) return rbm_log_ratio(d, x, theta, s, -1);
        int e = s - d->n_hx;
        return rbm_log_ratio(d, x, theta, d->edge_i[e], d->edge_j[e]);
    }
    if (d->guide != GFMC_GUIDE_JASTROW) return 0.0;
    if (s < d->n_hx) {
        double xk = (double)x[s];
        double mf = site_mf(d, x, s);
        return (-2.0 * xk) * mf - (2.0 * d->jas_h[s]) * xk;
    }
    int e = s - d->n_hx;
    int i = d->edge_i[e], j = d->edge_j[e];
    double xi = (double)x[i], xj = (double)x[j];
    double pair = xi * edge_mf(d, x, e, 0) + xj * edge_mf(d, x, e, 1);
    double hterm = d->jas_h[i] * xi + d->jas_h[j] * xj;
    return -2.0 * pair - 2.0 * hterm;
}

/* Bare SpinHalf off-diagonal element of slot s (SpinHalf.get_rates:316-369). */
static double slot_bare(const gfmc_problem_desc *d, const int8_t *x, int s) {
    if (s < d->n_hx) return d->hx_half[s];
    int e = s - d->n_hx;
    int anti = x[d->edge_i[e]] != x[d->edge_j[e]];
    return anti ? d->jpm[e] : d->jpmpm[e];
}

/* Guide-transformed rate of slot s: bare * exp(where(bare == 0, 0, ld))
 * (transition_rates, evolution.py:253-261). */
static double slot_rate(const gfmc_problem_desc *d, const int8_t *x, int s, double *ld_out,
                        const double *theta) {
    double bare = slot_bare(d, x, s);
    if (bare == 0.0 && !ld_out) return bare;
    double ld = slot_log_diff(d, x, s, theta);
    if (ld_out) *ld_out = ld;
    if (bare == 0.0) return bare;
    return bare * exp(ld);
}

static double *theta_of(const gfmc_problem_desc *d, const int8_t *x) {
    if (d->guide == GFMC_GUIDE_MLP) {
        double *t = (double *)malloc(sizeof(double) + sizeof(float) * mlp_grid_floats(d));
        mlp_grid_t g = mlp_grid_fill(d, (float *)(void *)(t + 1));
        t[0] = (double)mlp_log_psi_g(d, &g, x);
        return t;
    }
    if (d->guide != GFMC_GUIDE_RBM) return NULL;
    double *t = (double *)malloc(sizeof(double) * (size_t)d->rbm_hidden);
    orc_rbm_theta(d, x, t);
    return t;
}

/* Diagonal element H_xx (SpinHalf.diagonal_weight:291-314).  The non-uniform
 * sums use the blocked wave order. */
double orc_diag(const gfmc_problem_desc *d, const int8_t *x) {
    double diag = 0.0;
    if (d->has_jz) {
        if (d->uniform_jz) {
            int n_anti = 0;
            for (int e = 0; e < d->n_edges; ++e) n_anti += x[d->edge_i[e]] != x[d->edge_j[e]];
            int n_par = d->n_edges - n_anti;
            diag = (d->jz[0] / 4) * (double)(n_par - n_anti);
        } else {
            double *v = (double *)malloc(sizeof(double) * (size_t)(d->n_edges > 0 ? d->n_edges : 1));
            for (int e = 0; e < d->n_edges; ++e) {
                int anti = x[d->edge_i[e]] != x[d->edge_j[e]];
                v[e] = anti ? -d->jz[e] / 4.0 : d->jz[e] / 4.0;
            }
            diag = orc_blocked_scan(v, d->n_edges, NULL);
            free(v);
        }
    }
    if (d->has_hz) {
        double *v = (double *)malloc(sizeof(double) * (size_t)d->n_sites);
        for (int k = 0; k < d->n_sites; ++k) v[k] = d->hz[k] * (double)x[k];
        diag = diag + orc_blocked_scan(v, d->n_sites, NULL) / 2;
        free(v);
    }
    return diag;
}

/* Rates of one configuration in the reference layout [diag, off-diagonal
 * slots...] plus the guide log-ratios (slot 0: 0); returns e_L = diag - R
 * with R the blocked-order sum of the weights -rates[1:]. */
double orc_rates(const gfmc_problem_desc *d, const int8_t *x, double *rates, double *log_diff) {
    int n_off = n_off_of(d);
    double *w = (double *)malloc(sizeof(double) * (size_t)(n_off > 0 ? n_off : 1));
    double *theta = theta_of(d, x);
    double diag = orc_diag(d, x);
    if (rates) rates[0] = diag;
    if (log_diff) log_diff[0] = 0.0;
    for (int s = 0; s < n_off; ++s) {
        double ld;
        double r = slot_rate(d, x, s, &ld, theta);
        w[s] = -r;
        if (rates) rates[1 + s] = r;
        if (log_diff) log_diff[1 + s] = ld;
    }
    double R = orc_blocked_scan(w, n_off, NULL);
    free(w);
    free(theta);
    return diag - R;
}

/* log psi_T (JastrowSpinHalf.log_psi, jastrow_spin_half.py:299-324):
 * 0.5 * sum_k sum_q J_kq x_k x_other + sum_k h_k x_k, sequential order. */
double orc_log_psi(const gfmc_problem_desc *d, const int8_t *x) {
    if (d->guide == GFMC_GUIDE_MLP) return (double)orc_mlp_log_psi(d, x);
    if (d->guide == GFMC_GUIDE_RBM) {
        double *t = theta_of(d, x);
        double field = 0.0, hid = 0.0;
        for (int k = 0; k < d->n_sites; ++k) field += (double)d->rbm_a[k] * (double)x[k];
        for (int h = 0; h < d->rbm_hidden; ++h) hid += lncosh(t[h]);
        free(t);
        return field + hid;
    }
    if (d->guide != GFMC_GUIDE_JASTROW) return 0.0;
    int D = d->jas_max_deg;
    double pair = 0.0, field = 0.0;
    for (int k = 0; k < d->n_sites; ++k)
        for (int q = 0; q < D; ++q)
            pair += d->jas_site_J[k * D + q] * (double)x[k] * (double)x[d->jas_site_other[k * D + q]];
    for (int k = 0; k < d->n_sites; ++k) field += d->jas_h[k] * (double)x[k];
    return 0.5 * pair + field;
}

/* ------------------------------------------------------------------------ */
/* Spin packing                                                               */
/* ------------------------------------------------------------------------ */
static int n_words_of(int N) { return (N + 31) / 32; }

static void unpack1(const uint32_t *sp, int N, int8_t *x) {
    for (int k = 0; k < N; ++k) x[k] = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1 : -1;
}
static void pack1(const int8_t *x, int N, uint32_t *sp) {
    int W = n_words_of(N);
    for (int q = 0; q < W; ++q) sp[q] = 0u;
    for (int k = 0; k < N; ++k)
        if (x[k] > 0) sp[k >> 5] |= 1u << (k & 31);
}

void orc_unpack(int N, int64_t nw, const uint32_t *spins, double *config) {
    int W = n_words_of(N);
    for (int64_t w = 0; w < nw; ++w)
        for (int k = 0; k < N; ++k)
            config[w * N + k] = ((spins[w * W + (k >> 5)] >> (k & 31)) & 1u) ? 1.0 : -1.0;
}
void orc_pack(int N, int64_t nw, const double *config, uint32_t *spins) {
    int W = n_words_of(N);
    for (int64_t w = 0; w < nw; ++w) {
        for (int q = 0; q < W; ++q) spins[w * W + q] = 0u;
        for (int k = 0; k < N; ++k)
            if (config[w * N + k] > 0) spins[w * W + (k >> 5)] |= 1u << (k & 31);
    }
}

/* rates of a batch of packed walkers (reference layout, FP64) */
void orc_rates_batch(const gfmc_problem_desc *d, int64_t nw, const uint32_t *spins,
                     double *rates, double *e_loc, double *log_diff) {
    int N = d->n_sites, W = n_words_of(N), M = 1 + n_off_of(d);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t w = 0; w < nw; ++w) {
        int8_t *x = (int8_t *)malloc((size_t)N);
        unpack1(spins + w * W, N, x);
        double el = orc_rates(d, x, rates ? rates + w * M : NULL, log_diff ? log_diff + w * M : NULL);
        if (e_loc) e_loc[w] = el;
        free(x);
    }
}

void orc_log_psi_batch(const gfmc_problem_desc *d, int64_t nw, const uint32_t *spins, double *out) {
    int N = d->n_sites, W = n_words_of(N);
    for (int64_t w = 0; w < nw; ++w) {
        int8_t *x = (int8_t *)malloc((size_t)N);
        unpack1(spins + w * W, N, x);
        out[w] = orc_log_psi(d, x);
        free(x);
    }
}

/* ------------------------------------------------------------------------ */
/* Continuous-time propagation (evolve_walker_cont_time, evolution.py:559-721)*/
/* Reference semantics: every iteration re-evaluates every slot from the     */
/* current configuration (full re-enumeration, evolution.py:666-674); the    */
/* loop runs while time_used < beta (:697); the final e_loc is the e_L of    */
/* the last evaluation; log_w -= avg_log_weight * beta at the end (:718).    */
/* ------------------------------------------------------------------------ */
static void flip_slot(const gfmc_problem_desc *d, int8_t *x, int idx) {
    if (idx < d->n_hx) {
        x[idx] = (int8_t)-x[idx];
    } else {
        int e = idx - d->n_hx;
        x[d->edge_i[e]] = (int8_t)-x[d->edge_i[e]];
        x[d->edge_j[e]] = (int8_t)-x[d->edge_j[e]];
    }
}

/* every off-diagonal weight w[s] = -rate_s of x; returns the diagonal */
static double weights_of(const gfmc_problem_desc *d, const int8_t *x, double *w) {
    int n_off = n_off_of(d);
    double *theta = theta_of(d, x);
    for (int s = 0; s < n_off; ++s) w[s] = -slot_rate(d, x, s, NULL, theta);
    free(theta);
    return orc_diag(d, x);
}

/* method 0: continuous time (evolve_walker_cont_time, evolution.py:559-721);
 * 1: n_bra sub-steps of (lam - H) (evolve_walker_single_step, :322-424):
 *    stay with weight w0 = lam - H_xx, hop to slot s with weight w_s; the
 *    draw u = min(U (w0 + R), nextafter(w0 + R, 0)) stays when u < w0, else
 *    u - w0 is drawn from the blocked CDF of the slots; log_w += log(w0 + R);
 *    a total w0 + R <= 0 stays (the reference's sample_index returns slot
 *    M - 1 then, which is the stay when there are no off-diagonal slots;
 *    the walker's weight is zero either way);
 * 2: the same with geometric holds (evolve_walker_multi_step, :427-556):
 *    k = min(floor(log1p(-psi) / log1p(-R / b)) + 1, remaining + 1), the hop
 *    applied when k <= remaining, log_w += log(b) k, b = lam - H_xx + R.
 * The discrete methods end with one more evaluation for e_loc. */
static void propagate_one(const gfmc_problem_desc *d, int64_t widx, uint32_t *sp,
                          double *log_w, double *e_loc, double beta, double alw,
                          uint64_t seed, uint32_t step, uint32_t tag, int32_t *evals,
                          int8_t *x, double *w, int method, double lam) {
    int N = d->n_sites, n_off = n_off_of(d);
    unpack1(sp, N, x);
    double lw = *log_w, eL = 0.0;
    uint32_t it = 0;
    double incl[WAVE];
    if (method == 0) {
        double t = 0.0;
        while (t < beta) {
            double diag = weights_of(d, x, w);
            double R = orc_blocked_scan(w, n_off, incl);
            eL = diag - R;
            double U1, psi;
            draw2(seed, it, (uint32_t)widx, step, tag, &U1, &psi);
            int idx = -1;
            double tau_raw = INFINITY;
            /* a total that overflowed takes no hop (the kernels' hop_rate):
             * the walker holds for the rest of the step, its log weight
             * becomes non-finite; a zero hold time would never end the step */
            if (R > 0 && isfinite(R)) {
                double u = U1 * R, cap = nextafter(R, 0.0);
                if (u > cap) u = cap;
                idx = orc_sample_blocked(w, n_off, incl, u);
                tau_raw = orc_neg_log_unit(1.0 - psi) / R;
            }
            double rem = beta - t;
            double tau = tau_raw < rem ? tau_raw : rem;
            lw = lw - eL * tau;
            t = t + tau;
            ++it;
            if (tau_raw < rem && idx >= 0) flip_slot(d, x, idx);
        }
        lw = lw - alw * beta;
    } else {
        int n_bra = (int)beta;
        if (method == 1) {
            for (; it < (uint32_t)n_bra;) {
                double diag = weights_of(d, x, w);
                double R = orc_blocked_scan(w, n_off, incl);
                double w0 = lam - diag, W = w0 + R;
                double U1, psi;
                draw2(seed, it, (uint32_t)widx, step, tag, &U1, &psi);
                double u = U1 * W, cap = nextafter(W, 0.0);
                if (u > cap) u = cap;
                int idx = -1;
                /* a zero (or negative) total keeps the walker (it dies through
                 * log(W)): no slot has weight to draw */
                if (W > 0.0 && !(u < w0)) idx = orc_sample_blocked(w, n_off, incl, u - w0);
                lw = lw + log(w0 + R);
                ++it;
                if (idx >= 0) flip_slot(d, x, idx);
            }
        } else {
            int used = 0;
            while (used < n_bra) {
                double diag = weights_of(d, x, w);
                double R = orc_blocked_scan(w, n_off, incl);
                double bx = (lam - diag) + R;
                double U1, psi;
                draw2(seed, it, (uint32_t)widx, step, tag, &U1, &psi);
                int idx = -1;
                if (R > 0) {
                    double u = U1 * R, cap = nextafter(R, 0.0);
                    if (u > cap) u = cap;
                    idx = orc_sample_blocked(w, n_off, incl, u);
                }
                int remaining = n_bra - used;
                double p_hop = R / bx;
                double k_raw = (double)remaining + 1.0;
                if (p_hop > 0.0) {
                    double hold = floor(log1p(-psi) / log1p(-p_hop)) + 1.0;
                    k_raw = hold < k_raw ? hold : k_raw;
                }
                int kr = (int)k_raw;
                int k = kr < remaining ? kr : remaining;
                lw = lw + log(bx) * (double)k;
                used += k;
                ++it;
                if (kr <= remaining && idx >= 0) flip_slot(d, x, idx);
            }
        }
        double diag = weights_of(d, x, w);
        double R = orc_blocked_scan(w, n_off, incl);
        eL = diag - R;
        ++it;
        lw = lw - alw * (double)n_bra;
    }
    *log_w = lw;
    *e_loc = eL;
    if (evals) *evals = (int32_t)it;
    pack1(x, N, sp);
}

int orc_propagate(const gfmc_problem_desc *d, int64_t nw, uint32_t *spins, double *log_w,
                  double *e_loc, double beta, double alw, uint64_t seed, uint32_t step,
                  uint32_t tag, int32_t *evals, int n_threads, int method, double lam) {
    int N = d->n_sites, W = n_words_of(N), n_off = n_off_of(d);
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
#pragma omp parallel
    {
        int8_t *x = (int8_t *)malloc((size_t)N);
        double *w = (double *)malloc(sizeof(double) * (size_t)(n_off > 0 ? n_off : 1));
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < nw; ++i)
            propagate_one(d, i, spins + i * W, log_w + i, e_loc + i, beta, alw, seed, step, tag,
                          evals ? evals + i : NULL, x, w, method, lam);
        free(x);
        free(w);
    }
    return 0;
}

/* The same for a sample of a larger population: walker i of the buffers is
 * walker ids[i] of the population (its Philox counter index), so a few
 * walkers of a full-size GPU launch can be checked without propagating the
 * rest. */
int orc_propagate_ids(const gfmc_problem_desc *d, int64_t nw, const int64_t *ids,
                      uint32_t *spins, double *log_w, double *e_loc, double beta, double alw,
                      uint64_t seed, uint32_t step, uint32_t tag, int32_t *evals, int n_threads,
                      int method, double lam) {
    int N = d->n_sites, W = n_words_of(N), n_off = n_off_of(d);
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
#pragma omp parallel
    {
        int8_t *x = (int8_t *)malloc((size_t)N);
        double *w = (double *)malloc(sizeof(double) * (size_t)(n_off > 0 ? n_off : 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < nw; ++i)
            propagate_one(d, ids[i], spins + i * W, log_w + i, e_loc + i, beta, alw, seed, step,
                          tag, evals ? evals + i : NULL, x, w, method, lam);
        free(x);
        free(w);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Reconfiguration (reconfiguration.py:9-62) in block order: blocks of RB = */
/* 1024 walkers, 256 threads x 4 contiguous walkers; each thread adds its 4 */
/* walkers sequentially from 0, a block combines its 256 thread values with */
/* the wave tree inside each 64-wide wave and sequentially over the 4 wave  */
/* totals, and block totals are combined sequentially in block order.       */
/* ------------------------------------------------------------------------ */
#define RC_T 256
#define RC_PER 4
#define RB (RC_T * RC_PER)

static void rc_block_incl(const double *v, int64_t n, int64_t b, double *incl /*[RC_T]*/) {
    for (int t = 0; t < RC_T; ++t) {
        double acc = 0.0;
        for (int c = 0; c < RC_PER; ++c) {
            int64_t k = b * RB + (int64_t)t * RC_PER + c;
            if (k < n) acc += v[k];
        }
        incl[t] = acc;
    }
    for (int wv = 0; wv < RC_T / WAVE; ++wv) wave_tree(incl + wv * WAVE);
    double P = 0.0;
    for (int wv = 0; wv < RC_T / WAVE; ++wv) {
        double tot = incl[wv * WAVE + WAVE - 1];
        for (int l = 0; l < WAVE; ++l) incl[wv * WAVE + l] = P + incl[wv * WAVE + l];
        P = P + tot;
    }
}

/* block totals in block order, summed sequentially from 0 */
static double rc_total(const double *v, int64_t n, double *btot /* [B] or NULL */) {
    int64_t B = (n + RB - 1) / RB;
    double incl[RC_T], S = 0.0;
    for (int64_t b = 0; b < B; ++b) {
        rc_block_incl(v, n, b, incl);
        if (btot) btot[b] = incl[RC_T - 1];
        S = S + incl[RC_T - 1];
    }
    return S;
}

/* psi: the shared offset in [0, 1).  out2 <- {log G_n - log n, E_n}. */
int orc_reconfigure(int64_t n, const double *log_w, const double *e_loc, double psi,
                    int32_t *parents, double *out2, double *cdf_out) {
    int64_t B = (n + RB - 1) / RB;
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *wn = (double *)malloc(sizeof(double) * (size_t)n);
    double *cdf = (double *)malloc(sizeof(double) * (size_t)n);
    double *bw = (double *)malloc(sizeof(double) * (size_t)B);
    /* logsumexp with a non-finite max replaced by 0 (jax.scipy.special.logsumexp) */
    double m = -INFINITY;
    for (int64_t k = 0; k < n; ++k) m = log_w[k] > m ? log_w[k] : m;
    double m0 = isfinite(m) ? m : 0.0;
    for (int64_t k = 0; k < n; ++k) v[k] = exp(log_w[k] - m0);
    double logG = log(rc_total(v, n, NULL)) + m0;
    for (int64_t k = 0; k < n; ++k) wn[k] = exp(log_w[k] - logG);
    for (int64_t k = 0; k < n; ++k) v[k] = e_loc[k] * wn[k];
    double En = rc_total(v, n, NULL);
    /* prefix sum: block prefix + thread prefix, then sequential within the
     * thread's 4 walkers; then a running max so the CDF is non-decreasing */
    rc_total(wn, n, bw);
    double incl[RC_T], Pb = 0.0;
    for (int64_t b = 0; b < B; ++b) {
        rc_block_incl(wn, n, b, incl);
        for (int t = 0; t < RC_T; ++t) {
            double c = Pb + (t == 0 ? 0.0 : incl[t - 1]);
            for (int q = 0; q < RC_PER; ++q) {
                int64_t k = b * RB + (int64_t)t * RC_PER + q;
                if (k < n) {
                    c = c + wn[k];
                    cdf[k] = c;
                }
            }
        }
        Pb = Pb + bw[b];
    }
    for (int64_t k = 1; k < n; ++k)
        if (cdf[k] < cdf[k - 1]) cdf[k] = cdf[k - 1];
    int64_t last_pos = 0;
    for (int64_t k = 0; k < n; ++k)
        if (wn[k] > 0) last_pos = k;
    for (int64_t i = 0; i < n; ++i) {
        double u = ((double)i + psi) / (double)n;
        int64_t lo = 0, hi = n; /* first k with cdf[k] > u */
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (cdf[mid] > u) hi = mid; else lo = mid + 1;
        }
        parents[i] = (int32_t)(lo < last_pos ? lo : last_pos);
    }
    out2[0] = logG - log((double)n);
    out2[1] = En;
    if (cdf_out) memcpy(cdf_out, cdf, sizeof(double) * (size_t)n);
    free(v);
    free(wn);
    free(cdf);
    free(bw);
    return 0;
}

double orc_reconfig_psi(uint64_t seed, uint32_t step, uint32_t tag) {
    double a, b;
    draw2(seed, 0u, 0u, step, tag, &a, &b);
    return a;
}

void orc_gather(int N, int64_t n, const uint32_t *spins_in, const double *e_in,
                const int32_t *parents, uint32_t *spins_out, double *e_out) {
    int W = n_words_of(N);
    for (int64_t i = 0; i < n; ++i) {
        memcpy(spins_out + i * W, spins_in + (int64_t)parents[i] * W, sizeof(uint32_t) * (size_t)W);
        e_out[i] = e_in[parents[i]];
    }
}

/* ------------------------------------------------------------------------ */
/* Operator application (apply_operator_one, evolution.py:848-889).  The    */
/* operator's guide-transformed rates at x (slot 0 its diagonal element,    */
/* then the off-diagonal slots, as orc_rates lays them out); v = their      */
/* total in the blocked order over 64 lanes.  With v > 0 slot k is drawn by */
/* the blocked inverse CDF at u = min(U v, nextafter(v, 0)), U from Philox  */
/* (0, walker, step, tag), and the configuration moves by slot k - 1 when   */
/* k > 0; v <= 0 keeps the configuration with value 0 (the reference's      */
/* zero-weight guard); an operator without off-diagonal slots keeps it with */
/* its signed diagonal element.                                             */
/* ------------------------------------------------------------------------ */
void orc_apply_operator(const gfmc_problem_desc *d, int64_t nw, const uint32_t *spins,
                        uint64_t seed, uint32_t step, uint32_t tag, uint32_t *out,
                        double *vals) {
    int N = d->n_sites, W = n_words_of(N), M = 1 + n_off_of(d);
    int saved = g_lanes;
    g_lanes = WAVE;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int8_t *x = (int8_t *)malloc((size_t)N);
    double *rates = (double *)malloc(sizeof(double) * (size_t)M);
    for (int64_t w = 0; w < nw; ++w) {
        unpack1(spins + w * W, N, x);
        orc_rates(d, x, rates, NULL);
        double v;
        int idx = 0;
        if (M == 1) {
            v = rates[0];
        } else {
            double incl[WAVE];
            double R = orc_blocked_scan(rates, M, incl);
            if (R <= 0.0) {
                v = 0.0;
            } else {
                uint32_t ctr[4] = {0u, (uint32_t)w, step, tag};
                uint32_t r[4];
                orc_philox(ctr, key, r);
                double u = orc_u53(r[0], r[1]) * R;
                double cap = nextafter(R, 0.0);
                if (u > cap) u = cap;
                idx = orc_sample_blocked(rates, M, incl, u);
                v = R;
            }
        }
        if (idx > 0) flip_slot(d, x, idx - 1);
        pack1(x, N, out + w * W);
        vals[w] = v;
    }
    free(rates);
    free(x);
    g_lanes = saved;
}

/* ------------------------------------------------------------------------ */
/* Walker initialisers (walker.py:143-306), Philox-driven.                   */
/* ------------------------------------------------------------------------ */
/* Fisher-Yates from the top of an ascending base multiset (n_sites - n_up
 * down-spins, then n_up up-spins); swap partner of position i = floor(r *
 * (i + 1) / 2^32), r the (i % 4)-th word of Philox(i / 4, walker, 0, INIT). */
void orc_init_sz(int N, int64_t nw, int n_up, uint64_t seed, uint32_t *spins) {
    int W = n_words_of(N);
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int8_t *x = (int8_t *)malloc((size_t)N);
    for (int64_t w = 0; w < nw; ++w) {
        for (int k = 0; k < N; ++k) x[k] = k < N - n_up ? -1 : 1;
        uint32_t r[4];
        for (int i = N - 1; i >= 1; --i) {
            uint32_t ctr[4] = {(uint32_t)(i >> 2), (uint32_t)w, 0u, GFMC_TAG_INIT};
            orc_philox(ctr, key, r);
            uint32_t j = (uint32_t)(((uint64_t)r[i & 3] * (uint64_t)(i + 1)) >> 32);
            int8_t tmp = x[i];
            x[i] = x[j];
            x[j] = tmp;
        }
        pack1(x, N, spins + w * W);
    }
    free(x);
}

/* word q of walker w = word (q % 4) of Philox(q / 4, w, 1, INIT), masked to N. */
void orc_init_random(int N, int64_t nw, uint64_t seed, uint32_t *spins) {
    int W = n_words_of(N);
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t w = 0; w < nw; ++w) {
        for (int q = 0; q < W; ++q) {
            uint32_t ctr[4] = {(uint32_t)(q >> 2), (uint32_t)w, 1u, GFMC_TAG_INIT};
            uint32_t r[4];
            orc_philox(ctr, key, r);
            uint32_t v = r[q & 3];
            int rem = N - 32 * q;
            if (rem < 32) v &= (1u << rem) - 1u;
            spins[w * W + q] = v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* RBM training: local-energy variance minimisation (no reference           */
/* counterpart; the reference trains guides by NetKet VMC energy            */
/* minimisation, generic_netket_training.py:331-473).  Over n samples x_i:  */
/*   L = (1/n) sum_i (E_i - Ebar)^2,   E_i = sum_k rates_k(x_i)             */
/* and, by the definition of the rates (rates_k = H_k psi(x^k)/psi(x)),     */
/*   dL/dp = (2/n) sum_i d_i sum_{k>=1} rates_k(x_i) (O_p(x_i^k) - O_p(x_i)) */
/* with d_i = E_i - Ebar (clamped to +-clip*sqrt(L) when clip > 0), x^k the */
/* successor of slot k and O_p = d log psi / dp:                            */
/*   O_{W_hm} = tanh(theta_h) x_m,  O_{a_m} = x_m,  O_{b_h} = tanh(theta_h). */
/* Flat parameter order [W (H x N row-major), a (N), b (H)].  Every O is    */
/* computed from scratch for every successor, in FP64.                      */
/* ------------------------------------------------------------------------ */
static void rbm_O(const gfmc_problem_desc *d, const int8_t *x, double *theta, double *O) {
    int H = d->rbm_hidden, N = d->n_sites;
    orc_rbm_theta(d, x, theta);
    for (int h = 0; h < H; ++h) {
        double t = tanh(theta[h]);
        for (int m = 0; m < N; ++m) O[(size_t)h * N + m] = t * (double)x[m];
        O[(size_t)H * N + N + h] = t;
    }
    for (int m = 0; m < N; ++m) O[(size_t)H * N + m] = (double)x[m];
}

/* grad: H*N + N + H doubles; out2 = (mean E, variance L) */
void orc_rbm_vmin_grad(const gfmc_problem_desc *d, int64_t n, const uint32_t *spins, double clip,
                       double *grad, double *out2) {
    int N = d->n_sites, W = n_words_of(N), H = d->rbm_hidden, n_off = n_off_of(d);
    int M = 1 + n_off;
    size_t P = (size_t)H * N + N + H;
    double *rates = (double *)malloc(sizeof(double) * (size_t)n * M);
    double *E = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int8_t *x = (int8_t *)malloc((size_t)N), *xs = (int8_t *)malloc((size_t)N);
    double *theta = (double *)malloc(sizeof(double) * (size_t)H);
    double *O = (double *)malloc(sizeof(double) * P), *Os = (double *)malloc(sizeof(double) * P);
    for (size_t q = 0; q < P; ++q) grad[q] = 0.0;
    double mean = 0.0, var = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        unpack1(spins + i * W, N, x);
        E[i] = orc_rates(d, x, rates + i * M, NULL);
        mean += E[i];
    }
    mean /= (double)n;
    for (int64_t i = 0; i < n; ++i) var += (E[i] - mean) * (E[i] - mean);
    var /= (double)n;
    double lim = clip > 0.0 ? clip * sqrt(var) : INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        double dev = E[i] - mean;
        if (dev > lim) dev = lim;
        if (dev < -lim) dev = -lim;
        double g = 2.0 * dev / (double)n;
        unpack1(spins + i * W, N, x);
        rbm_O(d, x, theta, O);
        for (int s = 0; s < n_off; ++s) {
            double r = rates[i * M + 1 + s];
            if (r == 0.0) continue;
            memcpy(xs, x, (size_t)N);
            if (s < d->n_hx) {
                xs[s] = (int8_t)-xs[s];
            } else {
                int e = s - d->n_hx;
                xs[d->edge_i[e]] = (int8_t)-xs[d->edge_i[e]];
                xs[d->edge_j[e]] = (int8_t)-xs[d->edge_j[e]];
            }
            rbm_O(d, xs, theta, Os);
            for (size_t q = 0; q < P; ++q) grad[q] += g * r * (Os[q] - O[q]);
        }
    }
    out2[0] = mean;
    out2[1] = var;
    free(rates);
    free(E);
    free(x);
    free(xs);
    free(theta);
    free(O);
    free(Os);
}

/* theta (n x H) of a batch of packed configurations */
void orc_rbm_theta_batch(const gfmc_problem_desc *d, int64_t n, const uint32_t *spins,
                         double *theta) {
    int N = d->n_sites, W = n_words_of(N), H = d->rbm_hidden;
    int8_t *x = (int8_t *)malloc((size_t)N);
    for (int64_t i = 0; i < n; ++i) {
        unpack1(spins + i * W, N, x);
        orc_rbm_theta(d, x, theta + i * H);
    }
    free(x);
}

/* ------------------------------------------------------------------------ */
/* Diagonal operators of forward walking (driver.py:240-288): the value of  */
/* every column of a gfmc_diag_desc on packed walkers.  PAIR columns are    */
/* SpinHalf.diagonal_weight (SpinHalf.py:291-314): c0 * (n_par - n_anti)    */
/* when every edge coefficient is c0 (the reference's uniform form), else   */
/* the edge terms in edge order; then the site terms in site order.  SF     */
/* columns are SzSzStructureFactor.diagonal_weight (observables.py:162-165):*/
/* ((sum a s)^2 + (sum b s)^2) / 4 - c, sums in site order.                 */
/* ------------------------------------------------------------------------ */
static double diag_column(const gfmc_diag_desc *d, const int8_t *x, int v) {
    const int N = d->n_sites;
    const double *a = d->a + (int64_t)v * N;
    if (d->kind[v] == GFMC_DIAG_SF) {
        const double *b = d->b + (int64_t)v * N;
        double C = 0.0, S = 0.0;
        for (int k = 0; k < N; ++k) {
            C += x[k] > 0 ? a[k] : -a[k];
            S += x[k] > 0 ? b[k] : -b[k];
        }
        return 0.25 * (C * C + S * S) - d->c[v];
    }
    const int e0 = d->edge_off[v], e1 = d->edge_off[v + 1];
    int uniform = e1 > e0;
    for (int e = e0; e < e1; ++e)
        if (d->edge_c[e] != d->edge_c[e0]) uniform = 0;
    double val = 0.0;
    if (uniform) {
        int cnt = 0;
        for (int e = e0; e < e1; ++e) cnt += x[d->edge_i[e]] == x[d->edge_j[e]] ? 1 : -1;
        val = d->edge_c[e0] * (double)cnt;
    } else {
        for (int e = e0; e < e1; ++e)
            val += x[d->edge_i[e]] == x[d->edge_j[e]] ? d->edge_c[e] : -d->edge_c[e];
    }
    int linear = 0;
    for (int k = 0; k < N; ++k)
        if (a[k] != 0.0) linear = 1;
    if (linear) {
        double s = 0.0;
        for (int k = 0; k < N; ++k) s += x[k] > 0 ? a[k] : -a[k];
        val += s;
    }
    return val + d->c[v];
}

void orc_diag_values(const gfmc_diag_desc *d, int64_t nw, const uint32_t *spins, double *out) {
    const int N = d->n_sites, W = n_words_of(N), V = d->n_cols;
    int8_t *x = (int8_t *)malloc((size_t)N);
    for (int64_t i = 0; i < nw; ++i) {
        unpack1(spins + i * W, N, x);
        for (int v = 0; v < V; ++v) out[i * V + v] = diag_column(d, x, v);
    }
    free(x);
}

/* ------------------------------------------------------------------------ */
/* MLP guide training: O(x) = d log psi(x) / dp for the flat parameter      */
/* vector [W1, b1, W2, b2, w3, b3] by backpropagation in FP64, with the     */
/* straight-through BF16 rounding of h and the slopes of the tabulated tanh */
/* (256 D_j below saturation, 0 at it), and the variance gradient by its    */
/* definition, every O(x') from scratch (no reference counterpart).         */
/* ------------------------------------------------------------------------ */
static double mlp_slope_of(float z) {
    float t = fminf(fabsf(z) * 256.0f, 2047.9998779296875f);
    if (t == 2047.9998779296875f) return 0.0;
    mlp_table();
    return 256.0 * (double)mlp_D[(int)t];
}

static void mlp_O(const gfmc_problem_desc *d, const int8_t *x, double *O) {
    int N = d->n_sites, H1 = d->mlp_h1, H2 = d->mlp_h2;
    size_t oW1 = 0, ob1 = (size_t)H1 * N, oW2 = ob1 + H1, ob2 = oW2 + (size_t)H2 * H1;
    size_t ow3 = ob2 + H2, ob3 = ow3 + H2;
    float *h = (float *)malloc(sizeof(float) * (size_t)H1);
    double *s1 = (double *)malloc(sizeof(double) * (size_t)H1);
    double *d2 = (double *)malloc(sizeof(double) * (size_t)H2);
    for (int k = 0; k < H1; ++k) {
        double z = (double)orc_mlp_l1_grid(d->mlp_b1[k]);
        for (int n = 0; n < N; ++n)
            z += (double)x[n] * (double)orc_mlp_l1_grid(d->mlp_W1[(size_t)k * N + n]);
        h[k] = bf16_rne(orc_mlp_tanh((float)z));
        s1[k] = mlp_slope_of((float)z);
    }
    for (int c = 0; c < H2; ++c) {
        double z2 = 0.0;
        for (int k = 0; k < H1; ++k)
            z2 += (double)orc_bf16_flush(d->mlp_W2[(size_t)c * H1 + k]) * (double)h[k];
        float u = (float)z2 + d->mlp_b2[c];
        O[ow3 + c] = (double)orc_mlp_tanh(u);
        d2[c] = (double)d->mlp_w3[c] * mlp_slope_of(u);
        O[ob2 + c] = d2[c];
        for (int k = 0; k < H1; ++k) O[oW2 + (size_t)c * H1 + k] = d2[c] * (double)h[k];
    }
    O[ob3] = 1.0;
    for (int k = 0; k < H1; ++k) {
        double e = 0.0;
        for (int c = 0; c < H2; ++c)
            e += (double)orc_bf16_flush(d->mlp_W2[(size_t)c * H1 + k]) * d2[c];
        double d1 = e * s1[k];
        O[ob1 + k] = d1;
        for (int n = 0; n < N; ++n) O[oW1 + (size_t)k * N + n] = d1 * (double)x[n];
    }
    free(h);
    free(s1);
    free(d2);
}

void orc_mlp_O(const gfmc_problem_desc *d, const int8_t *x, double *O) { mlp_O(d, x, O); }

void orc_mlp_vmin_grad(const gfmc_problem_desc *d, int64_t n, const uint32_t *spins, double clip,
                       double *grad, double *out2) {
    int N = d->n_sites, W = n_words_of(N), n_off = n_off_of(d);
    int M = 1 + n_off, H1 = d->mlp_h1, H2 = d->mlp_h2;
    size_t P = (size_t)H1 * N + H1 + (size_t)H2 * H1 + 2 * (size_t)H2 + 1;
    double *rates = (double *)malloc(sizeof(double) * (size_t)n * M);
    double *E = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int8_t *x = (int8_t *)malloc((size_t)N), *xs = (int8_t *)malloc((size_t)N);
    double *O = (double *)malloc(sizeof(double) * P), *Os = (double *)malloc(sizeof(double) * P);
    for (size_t q = 0; q < P; ++q) grad[q] = 0.0;
    double mean = 0.0, var = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        unpack1(spins + i * W, N, x);
        E[i] = orc_rates(d, x, rates + i * M, NULL);
        mean += E[i];
    }
    mean /= (double)n;
    for (int64_t i = 0; i < n; ++i) var += (E[i] - mean) * (E[i] - mean);
    var /= (double)n;
    double lim = clip > 0.0 ? clip * sqrt(var) : INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        double dev = E[i] - mean;
        if (dev > lim) dev = lim;
        if (dev < -lim) dev = -lim;
        double g = 2.0 * dev / (double)n;
        unpack1(spins + i * W, N, x);
        mlp_O(d, x, O);
        for (int s = 0; s < n_off; ++s) {
            double r = rates[i * M + 1 + s];
            if (r == 0.0) continue;
            memcpy(xs, x, (size_t)N);
            if (s < d->n_hx) {
                xs[s] = (int8_t)-xs[s];
            } else {
                int e = s - d->n_hx;
                xs[d->edge_i[e]] = (int8_t)-xs[d->edge_i[e]];
                xs[d->edge_j[e]] = (int8_t)-xs[d->edge_j[e]];
            }
            mlp_O(d, xs, Os);
            for (size_t q = 0; q < P; ++q) grad[q] += g * r * (Os[q] - O[q]);
        }
    }
    out2[0] = mean;
    out2[1] = var;
    free(rates);
    free(E);
    free(x);
    free(xs);
    free(O);
    free(Os);
}
