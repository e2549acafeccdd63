// gfmc_hip.hip -- continuous-time GFMC on MI355X (gfx950): walker
// propagation, reconfiguration and their helpers behind the C-ABI of
// include/gfmc.h.
//
// Execution model: one 64-lane wavefront per walker.  The walker's spins
// (one byte per site), its off-diagonal weights and the edge states live in
// LDS for the whole branching step; HBM holds only the bit-packed spins and
// the FP64 per-walker scalars.  After a move only the slots whose rate
// depends on a flipped spin are re-evaluated (host-built dependency lists);
// every re-evaluated slot is computed from scratch, so it is bit-identical to
// a full re-enumeration (the reference evaluates every slot at every hop,
// evolution.py:666-674).
//
// Summation order of the off-diagonal weights ("blocked wave order"): slot s
// belongs to lane s / K with K = ceil(n_off / 64); each lane adds its block
// sequentially from 0, and the lane sums are combined by a Hillis-Steele
// inclusive scan.  LDS stores slot s at (s % K) * 64 + s / K so the lane
// loops are bank-conflict free.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gfmc.h"
#include "gfmc_math.h"
#include "philox.h"

#define WAVE 64
#define RECONF_THREADS 1024

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
// device counter added to the step of every launch this thread enqueues
// while a step loop is captured into a graph (null otherwise)
static thread_local const uint32_t *g_step_base = nullptr;

// Carried slot weights of the record path (driver.py's carried rates,
// initial_carried / gather_carried): a propagation ends by storing every
// walker's slot weights and anti-parallel count; the next one, after
// resampling, starts walker w from the rows of its parent instead of
// re-evaluating every slot.  The stored weights are the exact values a full
// evaluation of the same configuration gives (each weight is always
// eval_rec of the current configuration), so results are bit-identical.
// state[0] = rows valid, state[1] = buffer holding the latest rows,
// state[2] = lanes per walker they were laid out for (the tied-RBM kernel:
// its own tag), state[3] = the tied RBM's parameter generation.  The tied-RBM
// rows hold the slot weights, theta (exact in FP64) and the final spins; f_off
// is then theta's offset.
struct CarryDev {
    void *w[2];
    int32_t *anti[2];
    int32_t *state;
    const int32_t *parents;
    int64_t row;     // elements per walker row
    int64_t f_off;   // element offset of the local fields in a row (0: none)
    int64_t s_off;   // element offset of the walker's final spins in a row
    int on;
};
static thread_local const CarryDev *g_carry = nullptr;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(GFMC_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------------------
// device problem
// ---------------------------------------------------------------------------
template <typename T>
struct Tables {
    const T *jpm, *jpmpm, *hx_half, *jas_h, *jas_site_J, *jas_edge_J;
    // uniform values used by the record path
    T jpm0, jpmpm0, hx0, J0, h0;
    // its local-field form: log2 of a Jastrow ratio = lJ * (integer pair sum)
    // + lh * (integer spin sum), lJ = -2 J0 log2(e), lh = -2 h0 log2(e)
    T lJ, lh;
};

// Record path ("fast"): usable when the couplings are uniform, the Jastrow
// coupling is tied (every J is J0 or a masked 0), the field h is uniform,
// every site has at most REC_D partners and N, n_off < 32768.  Each slot is
// described by one 32-byte record of 16 int16 fields:
//   [0] slot  [1] site i  [2] site j (-1 for a site flip)  [3] partner mask
//   [4 .. 4+2D) partners: D of site i (excluding j), then D of site j
// and the slots a move s re-evaluates are stored as whole records at
// mv_rec[s * R + t] (R records per move, slot -1 = padding), so a move needs
// one round trip to HBM/L2.  For uniform J_z the anti-parallel count is
// updated from mv_nbr[s * RN + t] = neighbour p | (flipped site f << 16) of
// the flipped sites (edges between two flipped sites excluded).
#define REC_D 6

struct DevProblem {
    int N, n_edges, n_hx, n_jpm, n_off;
    int has_jz, uniform_jz, has_hz, guide;
    int words, K, Ke, Ks, max_deg, full;
    int n_pad, e_pad;
    double jz0;
    const int *edge_i, *edge_j;
    const double *jz, *hz;
    const int *jas_site_other, *jas_edge_other;
    const int *dep_off, *dep_slots, *edge_off, *dep_edges;
    // record path
    int fast, R, RN, RE, flags;
    int method;       // GFMC_METHOD_* of the propagation kernels
    int wpb, xcd;     // k_propagate launch shape (gfmc_problem_set_launch)
    double lam;       // the shift lambda of the discrete methods
    const uint4 *slot_rec;   // [K * 64 * 2], in LDS order (record p -> LDS position p)
    // the same per walker-group size G = 64, 32, 16: K_G = ceil(n_off / G),
    // records in that group's LDS order, and whether the record path fits G
    // lanes (every move's records and neighbours within G)
    const uint4 *slot_rec_g[3];
    int Kg[3], g_ok[3];
    const uint4 *mv_rec;     // [n_off * R * 2]
    const int *mv_nbr;       // [n_off * RN]
    const int *mv_sites;     // [n_off]  i | (j << 16), j = 0xffff for a site flip
    const int *mv_edges;     // [n_off * RE] edges whose state a move may change (-1 pad)
    const int *edge_pack;    // [n_edges] i | (j << 16)
    // local-field form of the records (FP32 kernels, when fld): records of two
    // words {slot | i << 16, j | flags << 16} in the LDS order of each group
    // size, and per move and group size {LDS position | i << 16, ...}; the
    // field updates of each move, mv_fnb[s * FD + t] =
    // partner | (flipped site << 16), t < D for site i and D <= t < 2D for
    // site j; every site's partners, 8 int16 per site; site N, a zero spin,
    // as padding where an entry pads (zsite: irregular partner counts, or
    // site flips with pair moves); f_pad bytes of fields per walker
    int fld, FD, f_pad, zsite;
    const uint2 *slot_rf_g[3];
    const uint2 *mv_rf_g[3]; // [n_off * R] per group layout, first field the LDS position
    const int *mv_fnb;       // [n_off * FD]
    const uint4 *fsite;      // [N]
    // RBM guide (FP32 arithmetic on BF16-valued parameters)
    int H, w_words, rbm_wave_bytes;
    const uint32_t *rbm_W;   // [H][N] BF16, packed two per word (N even-padded)
    const float *rbm_a, *rbm_b, *rbm_cs;   // [N], [H], column sums sum_h W_hk [N]
    // lanes-over-hidden-units kernels (gfmc_rbm.inl): rbh = their M = Hp / 64
    // (0: the hidden layer is too wide for them), WT [N][Hp] BF16 bits,
    // the per-launch walker queue
    int rbh, Hp, wt_bytes, rbh_wave_bytes;
    const uint16_t *rbm_WT;
    int *rbm_queue;
    // translation-tied RBM on a torus (gfmc_rbm_set_tied; gfmc_rbm.inl): rbt
    // = 1 when on, with rbt_cp pair and rbt_cs single lane chunks (rbt_pk:
    // one single chunk of at most 16 units, packed); raw
    // tables rbt_raw ((2 rbt_ncls + 1) rbt_F blocks of rbt_Q, then rbt_Q
    // zeros) and the kernel's table rbt_tab of rbt_words floats (entry idx =
    // (raw[idx], raw[idx - rbt_D]), rounded up to a multiple of 4 floats)
    int rbt, rbt_cp, rbt_cs, rbt_pk, rbt_D, rbt_F, rbt_nT, rbt_Q, rbt_ncls, rbt_words;
    int rbt_gen;   // parameter generation: carried rows of older ones are not used
    const float *rbt_tab;
    float *rbt_raw;
    const int *rbt_unit;    // [(cp + cs) 64] off(a) of each lane chunk, padding F Q + O
    const int *rbt_hid;     // [(2 cp + cs) 64] hidden unit of each lane register, -1: pad
    const int *rbt_scode;   // [n_off]  code(first) | class << 20 | swapped << 30
    const int *rbt_kcode;   // [N]      code(k)
    const int *rbt_qsite;   // [Q]      the site of entry q
    const int *rbt_qpair;   // [ncls Q] the partner site of entry q in class c, -1: none
    // MLP guide (layer-1 sums exact integers at scale 2^16, layer 2 on MFMA);
    // rbm_wave_bytes is also the per-wavefront LDS of its walker kernels
    int mlp_h1, mlp_h2;
    const int32_t *mlp_W1I;   // [N][h1] layer-1 grid values x 2^16 (site-major)
    const double *mlp_b1;     // [h1] layer-1 grid values
    const uint16_t *mlp_W2;   // [h2][h1] BF16 bits
    const float *mlp_b2, *mlp_w3;
    const float *mlp_b3p;     // [1]
    const float2 *mlp_tab;    // [2048] {tanh(j / 256), tanh((j + 1) / 256) - tanh(j / 256)}
};

struct MlpWork;
struct MlpTrainWork;
static void free_mlp_work(MlpWork *m);
static void free_mlp_train(MlpTrainWork *t);
static MlpWork *new_mlp_work(int blocks_per_walker);

struct gfmc_problem {
    DevProblem dp;
    MlpWork *mlp = nullptr;   // the MLP guide's propagation workspace (grown on demand)
    mutable MlpTrainWork *mlpt = nullptr;   // its training workspace (grown on demand)
    Tables<double> t64;
    Tables<float> t32;
    int precision;
    std::vector<void *> allocs;
    int64_t lds_bytes;
    int64_t rbh_lds_bytes = 0;   // LDS per workgroup of the gfmc_rbm.inl kernels
    int64_t rbt_lds_bytes = 0;   // LDS per workgroup of the tied-RBM kernel
    int rbt_waves = 0;           // its wavefronts per workgroup
    int jpmpm_zero = 0;          // every edge's J_pmpm is zero
    int n_cu;
    int lanes = WAVE;   // lanes per walker of the propagation kernel (64, 32, 16)
};

static int gidx_of(int lanes) { return lanes == 64 ? 0 : (lanes == 32 ? 1 : 2); }

static int block_k_of(int n) { return n > 0 ? (n + WAVE - 1) / WAVE : 1; }
static int round_up(int n, int m) { return (n + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
// DPP lane moves: value of the lane the control selects, 0 where that lane
// is outside the row (bound_ctrl) or this lane's row is masked off.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_get(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xf, true));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_get(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Inclusive scan over the 64 lanes in "wave tree" order: Hillis-Steele
// inside each 16-lane row (row_shr 1, 2, 4, 8), then rows 1 and 3 add the
// last lane of the row before (row_bcast:15), then rows 2 and 3 add lane 31
// (row_bcast:31).  All data movement is DPP; no LDS round trips.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v, int lane) {
    (void)lane;
    v += dpp_get<0x111, 0xf>(v);
    v += dpp_get<0x112, 0xf>(v);
    v += dpp_get<0x114, 0xf>(v);
    v += dpp_get<0x118, 0xf>(v);
    v += dpp_get<0x142, 0xa>(v);
    v += dpp_get<0x143, 0xc>(v);
    return v;
}

__device__ __forceinline__ float lane_bcast(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int lane_bcast(int v, int src) {
    return __builtin_amdgcn_readlane(v, src);
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

// Walker groups: G = 64, 32 or 16 lanes per walker (64 / G walkers per
// wavefront).  A group is G consecutive lanes; its scan is the first stages
// of the wave tree (the 16-lane rows, then row_bcast:15 for G >= 32, then
// row_bcast:31 for G = 64), so G = 64 is the wave tree itself.
template <int G, typename T>
__device__ __forceinline__ T grp_incl_scan(T v) {
    v += dpp_get<0x111, 0xf>(v);
    v += dpp_get<0x112, 0xf>(v);
    v += dpp_get<0x114, 0xf>(v);
    v += dpp_get<0x118, 0xf>(v);
    if constexpr (G >= 32) v += dpp_get<0x142, 0xa>(v);
    if constexpr (G == 64) v += dpp_get<0x143, 0xc>(v);
    return v;
}

// value of group lane src (group-uniform src)
template <int G, typename T>
__device__ __forceinline__ T grp_bcast(T v, int src) {
    if constexpr (G == WAVE) return lane_bcast(v, src);
    else return __shfl(v, src, G);
}

template <int G>
__device__ __forceinline__ int grp_sum_int(int v) {
#pragma unroll
    for (int d = G / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, G);
    return v;
}

// ballot of the group, bit g = group lane g
template <int G>
__device__ __forceinline__ unsigned long long grp_ballot(bool pred, int lane) {
    const unsigned long long b = __ballot(pred);
    if constexpr (G == WAVE) {
        (void)lane;
        return b;
    } else {
        return (b >> (lane & ~(G - 1))) & ((1ull << G) - 1ull);
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// index of G in the per-group tables (64 -> 0, 32 -> 1, 16 -> 2)
template <int G>
constexpr int gidx() { return G == 64 ? 0 : (G == 32 ? 1 : 2); }

__device__ __forceinline__ double dev_exp(double v) { return exp(v); }
__device__ __forceinline__ float dev_exp(float v) { return expf(v); }

// next representable value towards 0 of a positive finite number
__device__ __forceinline__ double below(double r) {
    return __longlong_as_double(__double_as_longlong(r) - 1);
}
__device__ __forceinline__ float below(float r) { return __int_as_float(__float_as_int(r) - 1); }

// a total rate that starts a hop: positive and finite.  A total that
// overflowed (a guide ratio beyond the range of T) takes no hop: the walker
// holds for the rest of the step, its log weight becomes non-finite (-e_L
// = +inf times the remaining time) and the driver's population check
// raises; a hop would have a zero hold time and the step would never end.
// NaN fails the comparisons as well.
template <typename T>
__device__ __forceinline__ bool hop_rate(T R) {
    return R > (T)0 && __builtin_isfinite(R);
}

template <typename T>
__device__ __forceinline__ T uniform_for(gfmc_u4 r);
template <>
__device__ __forceinline__ double uniform_for<double>(gfmc_u4 r) {
    return gfmc_u53(r.x, r.y);
}
template <>
__device__ __forceinline__ float uniform_for<float>(gfmc_u4 r) {
    return gfmc_u24(r.x);
}

// ---------------------------------------------------------------------------
// rates of one slot (SpinHalf.get_rates + JastrowSpinHalf ratios)
// ---------------------------------------------------------------------------
// Guide log-ratio of slot s (jastrow_spin_half.py:290-297, 462-472).
template <typename T>
__device__ __forceinline__ T slot_log_diff(const DevProblem &P, const Tables<T> &tb,
                                           const signed char *x, int s) {
    const int D = P.max_deg;
    if (s < P.n_hx) {
        T mf = (T)0;
        for (int q = 0; q < D; ++q)
            mf += tb.jas_site_J[s * D + q] * (T)x[P.jas_site_other[s * D + q]];
        T xk = (T)x[s];
        return ((T)-2 * xk) * mf - ((T)2 * tb.jas_h[s]) * xk;
    }
    const int e = s - P.n_hx;
    const int i = P.edge_i[e], j = P.edge_j[e];
    const int *oi = P.jas_edge_other + (size_t)e * 2 * D;
    const T *Ji = tb.jas_edge_J + (size_t)e * 2 * D;
    T mfi = (T)0, mfj = (T)0;
    for (int q = 0; q < D; ++q) mfi += Ji[q] * (T)x[oi[q]];
    for (int q = 0; q < D; ++q) mfj += Ji[D + q] * (T)x[oi[D + q]];
    T xi = (T)x[i], xj = (T)x[j];
    T pair = xi * mfi + xj * mfj;
    T hterm = tb.jas_h[i] * xi + tb.jas_h[j] * xj;
    return (T)-2 * pair - (T)2 * hterm;
}

template <typename T>
__device__ __forceinline__ T slot_bare(const DevProblem &P, const Tables<T> &tb,
                                       const signed char *x, int s) {
    if (s < P.n_hx) return tb.hx_half[s];
    const int e = s - P.n_hx;
    return (x[P.edge_i[e]] != x[P.edge_j[e]]) ? tb.jpm[e] : tb.jpmpm[e];
}

// weight = -rate = -(bare * exp(where(bare == 0, 0, log_diff)))
// (transition_rates, evolution.py:253-261)
template <typename T>
__device__ __forceinline__ T slot_weight(const DevProblem &P, const Tables<T> &tb,
                                         const signed char *x, int s) {
    T bare = slot_bare(P, tb, x, s);
    if (bare == (T)0 || P.guide != GFMC_GUIDE_JASTROW) return -bare;
    return -(bare * dev_exp(slot_log_diff(P, tb, x, s)));
}

// ---------------------------------------------------------------------------
// record path: the same slot arithmetic from a pre-joined 32-byte record
// ---------------------------------------------------------------------------
struct Rec {
    uint32_t w[8];
};

__device__ __forceinline__ Rec load_rec(const uint4 *base, size_t r) {
    const uint4 a = base[2 * r], b = base[2 * r + 1];
    Rec o;
    o.w[0] = a.x; o.w[1] = a.y; o.w[2] = a.z; o.w[3] = a.w;
    o.w[4] = b.x; o.w[5] = b.y; o.w[6] = b.z; o.w[7] = b.w;
    return o;
}

__device__ __forceinline__ Rec empty_rec() {
    Rec o;
#pragma unroll
    for (int q = 0; q < 8; ++q) o.w[q] = 0xffffffffu;
    return o;
}

// int16 field k (sign-extended); k must be a compile-time constant
__device__ __forceinline__ int rec_f(const Rec &r, int k) {
    int v = (int)((r.w[k >> 1] >> ((k & 1) * 16)) & 0xffffu);
    return (v ^ 0x8000) - 0x8000;
}

// slot weight from a record: identical operations, in the same order, as
// slot_weight on the generic tables
template <typename T>
__device__ __forceinline__ T eval_rec(const DevProblem &P, const Tables<T> &tb, const Rec &r,
                                      const signed char *x) {
    const int i = rec_f(r, 1), j = rec_f(r, 2), mask = rec_f(r, 3) & 0xffff;
    const bool jas = P.guide == GFMC_GUIDE_JASTROW;
    // all REC_D partner fields are read: the fields past the guide's degree
    // hold site i with a clear mask bit, so their terms are +-0 and leave
    // the sums bit-identical to sums over the degree; no branches, so the
    // partner spins are read from LDS in one batch
    if (j < 0) {
        const T bare = tb.hx0;
        if (bare == (T)0 || !jas) return -bare;
        signed char xs[REC_D];
#pragma unroll
        for (int q = 0; q < REC_D; ++q) xs[q] = x[rec_f(r, 4 + q)];
        T mf = (T)0;
#pragma unroll
        for (int q = 0; q < REC_D; ++q) mf += (((mask >> q) & 1) ? tb.J0 : (T)0) * (T)xs[q];
        const T xk = (T)x[i];
        return -(bare * dev_exp(((T)-2 * xk) * mf - ((T)2 * tb.h0) * xk));
    }
    const T xi = (T)x[i], xj = (T)x[j];
    const T bare = (x[i] != x[j]) ? tb.jpm0 : tb.jpmpm0;
    if (bare == (T)0 || !jas) return -bare;
    signed char xs[2 * REC_D];
#pragma unroll
    for (int q = 0; q < 2 * REC_D; ++q) xs[q] = x[rec_f(r, 4 + q)];
    T mfi = (T)0, mfj = (T)0;
#pragma unroll
    for (int q = 0; q < REC_D; ++q) mfi += (((mask >> q) & 1) ? tb.J0 : (T)0) * (T)xs[q];
#pragma unroll
    for (int q = 0; q < REC_D; ++q)
        mfj += (((mask >> (REC_D + q)) & 1) ? tb.J0 : (T)0) * (T)xs[REC_D + q];
    const T pair = xi * mfi + xj * mfj;
    const T hterm = tb.h0 * xi + tb.h0 * xj;
    return -(bare * dev_exp((T)-2 * pair - (T)2 * hterm));
}

// Local-field form of the FP32 record path.  F_k, the sum of the spins of
// site k's Jastrow partners, is kept in LDS (one byte per site) and updated
// by every move, so a slot weight reads two spins and two fields instead of
// 2 REC_D partner spins.  The host enables it when every record's partners
// are the Jastrow partners of its site(s), less the other flipped site where
// that one is a partner (flag bits 0 and 1); then the partner sums of
// eval_rec are s_i = F_i - c_i x_j and s_j = F_j - c_j x_i, exact integers,
// and the weight is -bare 2^(lJ (x_i s_i + x_j s_j) + lh (x_i + x_j)) (a
// site flip: -hx0 2^(lJ x_k F_k + lh x_k)); the same value as eval_rec up
// to FP32 rounding of the exponent.
__device__ __forceinline__ float eval_recf(const Tables<float> &tb, uint2 r,
                                           const signed char *x, const signed char *F) {
    const int i = (int)(r.x >> 16);
    const int j = (int)(int16_t)(r.y & 0xffffu);
    const int fl = (int)(r.y >> 16);
    if (j < 0) {
        const int xk = x[i], fk = F[i];
        const float e = tb.lJ * (float)(xk * fk) + tb.lh * (float)xk;
        return tb.hx0 == 0.0f ? -tb.hx0 : -(tb.hx0 * __builtin_amdgcn_exp2f(e));
    }
    const int xi = x[i], xj = x[j], fi = F[i], fj = F[j];
    const float bare = xi != xj ? tb.jpm0 : tb.jpmpm0;
    const int si = fi - (fl & 1) * xj, sj = fj - ((fl >> 1) & 1) * xi;
    const float e = tb.lJ * (float)(xi * si + xj * sj) + tb.lh * (float)(xi + xj);
    return bare == 0.0f ? -bare : -(bare * __builtin_amdgcn_exp2f(e));
}

// the local fields of the walker's spins (spins already in LDS); site N is
// the zero spin that padding entries of the partner tables name (present
// when zsite: irregular partner counts, or site flips with pair moves)
template <int G>
__device__ __forceinline__ void fld_init(const DevProblem &P, signed char *x, signed char *F,
                                         int gl) {
    if (P.zsite && gl == 0) x[P.N] = 0;
    wave_sync();
    const int D = P.max_deg;
    for (int k = gl; k < P.N; k += G) {
        const uint4 q = P.fsite[k];
        const uint32_t f[4] = {q.x, q.y, q.z, q.w};
        int s = 0;
        for (int t = 0; t < D; ++t) {
            const uint32_t w = f[t >> 1];
            s += x[(t & 1) ? (w >> 16) : (w & 0xffffu)];
        }
        F[k] = (signed char)s;
    }
    wave_sync();
}

// Diagonal element (SpinHalf.diagonal_weight, SpinHalf.py:291-314); gl is
// the lane in the walker's group of G lanes (blocked order over G lanes).
template <int G = WAVE>
__device__ double diag_value(const DevProblem &P, const signed char *x,
                             const unsigned char *anti, int n_anti, int gl) {
    double d = 0.0;
    if (P.has_jz) {
        if (P.uniform_jz) {
            d = (P.jz0 / 4) * (double)((P.n_edges - n_anti) - n_anti);
        } else {
            const int Ke = (P.n_edges + G - 1) / G;
            double acc = 0.0;
            for (int k = 0; k < Ke; ++k) {
                int e = gl * Ke + k;
                if (e < P.n_edges) acc += anti[e] ? -P.jz[e] / 4.0 : P.jz[e] / 4.0;
            }
            acc = grp_incl_scan<G>(acc);
            d = grp_bcast<G>(acc, G - 1);
        }
    }
    if (P.has_hz) {
        const int Ks = (P.N + G - 1) / G;
        double acc = 0.0;
        for (int k = 0; k < Ks; ++k) {
            int s = gl * Ks + k;
            if (s < P.N) acc += P.hz[s] * (double)x[s];
        }
        acc = grp_incl_scan<G>(acc);
        d = d + grp_bcast<G>(acc, G - 1) / 2;
    }
    return d;
}

// ---------------------------------------------------------------------------
// per-walker LDS state
// ---------------------------------------------------------------------------
template <typename T>
struct WalkerLds {
    T *w;                 // [K*64], slot s at (s % K) * 64 + s / K
    signed char *x;       // [n_pad]
    unsigned char *anti;  // [e_pad]
    signed char *f;       // [f_pad] local fields (the records' local-field form)
};

template <typename T>
__device__ __forceinline__ WalkerLds<T> carve(const DevProblem &P, unsigned char *lds) {
    WalkerLds<T> L;
    L.w = (T *)lds;
    L.x = (signed char *)(lds + (size_t)P.K * WAVE * sizeof(T));
    L.anti = (unsigned char *)(L.x + P.n_pad);
    L.f = (signed char *)(L.anti + P.e_pad);
    return L;
}

// LDS bytes of one walker in a group of G lanes, and its carve
template <typename T, int G>
__host__ __device__ inline int walker_lds_bytes(const DevProblem &P) {
    const int K = (P.n_off > 0 ? P.n_off + G - 1 : G) / G;
    return (int)(((size_t)K * G * sizeof(T) + P.n_pad + P.e_pad + P.f_pad + 15) / 16 * 16);
}

template <typename T, int G>
__device__ __forceinline__ WalkerLds<T> carve_g(const DevProblem &P, unsigned char *lds) {
    WalkerLds<T> L;
    L.w = (T *)lds;
    L.x = (signed char *)(lds + (size_t)P.Kg[gidx<G>()] * G * sizeof(T));
    L.anti = (unsigned char *)(L.x + P.n_pad);
    L.f = (signed char *)(L.anti + P.e_pad);
    return L;
}

__device__ __forceinline__ int wpos(int s, int K) { return (s % K) * WAVE + s / K; }

template <int G>
__device__ __forceinline__ int wpos_g(int s, int K) { return (s % K) * G + s / K; }

// Spins between the packed words (bit k of word k / 32: site k up) and LDS
// bytes (+1 / -1), four sites per lane operation: bits b0..b3 of a nibble
// spread to bytes by n * 0x00204081 & 0x01010101 (bit q lands at 8q, no
// carries); a byte 0 or 1 becomes -1 or +1 as 0x01 | (~(b * 0xff) & 0xfe).
// The bytes of a partial last quad lie inside the n_pad bytes of x.
template <int G = WAVE>
__device__ __forceinline__ void load_spins(const DevProblem &P, const uint32_t *sp,
                                           signed char *x, int gl) {
    const int nq = (P.N + 3) >> 2;
    for (int t = gl; t < nq; t += G) {
        const uint32_t nib = (sp[t >> 3] >> ((t & 7) * 4)) & 0xfu;
        const uint32_t b = (nib * 0x00204081u) & 0x01010101u;
        *(uint32_t *)(x + 4 * t) = 0x01010101u | (~(b * 0xffu) & 0xfefefefeu);
    }
}

// The inverse: bit 1 of a byte's complement is 1 for +1 (0x01) and 0 for
// -1 (0xff); (v * 0x00204081) >> 21 gathers the four byte bits to a nibble;
// the eight nibbles of a word are OR-combined across eight consecutive
// lanes.  Bits past site N - 1 stay clear.
template <int G = WAVE>
__device__ __forceinline__ void store_spins(const DevProblem &P, uint32_t *sp,
                                            const signed char *x, int gl,
                                            uint32_t *copy = nullptr) {
    const int nq = (P.N + 3) >> 2;
    for (int t0 = 0; t0 < P.words * 8; t0 += G) {
        const int t = t0 + gl;
        uint32_t nib = 0u;
        if (t < nq) {
            const uint32_t d = *(const uint32_t *)(x + 4 * t);
            const uint32_t v = (~d >> 1) & 0x01010101u;
            nib = ((v * 0x00204081u) >> 21) & 0xfu;
            const int left = P.N - 4 * t;
            if (left < 4) nib &= (1u << left) - 1u;
        }
        uint32_t w = nib << ((t & 7) * 4);
        w |= (uint32_t)__shfl_xor((int)w, 1, G);
        w |= (uint32_t)__shfl_xor((int)w, 2, G);
        w |= (uint32_t)__shfl_xor((int)w, 4, G);
        if ((t & 7) == 0 && t < P.words * 8) {
            sp[t >> 3] = w;
            if (copy) copy[t >> 3] = w;
        }
    }
}

// every slot weight, every edge state; returns the anti-parallel count.
// gl: lane in the walker's group of G lanes (the generic path: G = 64).
template <typename T, int FAST, int G = WAVE>
__device__ int full_eval(const DevProblem &P, const Tables<T> &tb, WalkerLds<T> &L, int gl) {
    if constexpr (FAST == 2) {
        // local-field records (fields already initialised), four in flight
        const int n = P.Kg[gidx<G>()] * G;
        const uint2 *recs = P.slot_rf_g[gidx<G>()];
        for (int p0 = 0; p0 < n; p0 += 4 * G) {
            uint2 r[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int p = p0 + q * G + gl;
                r[q] = p < n ? recs[p] : make_uint2(0xffffffffu, 0xffffffffu);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int p = p0 + q * G + gl;
                if (p < n)
                    L.w[p] = (int16_t)(r[q].x & 0xffffu) >= 0 ? eval_recf(tb, r[q], L.x, L.f)
                                                             : (T)0;
            }
        }
    } else if constexpr (FAST) {
        // slot records are stored in LDS order (record p describes the slot
        // kept at LDS position p, slot -1 = padding): coalesced reads,
        // conflict-free writes, four records in flight per lane
        const int n = P.Kg[gidx<G>()] * G;
        const uint4 *recs = P.slot_rec_g[gidx<G>()];
        for (int p0 = 0; p0 < n; p0 += 4 * G) {
            Rec r[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int p = p0 + q * G + gl;
                r[q] = p < n ? load_rec(recs, p) : empty_rec();
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int p = p0 + q * G + gl;
                if (p < n) L.w[p] = rec_f(r[q], 0) >= 0 ? eval_rec(P, tb, r[q], L.x) : (T)0;
            }
        }
    } else {
        static_assert(FAST || G == WAVE, "the generic path runs one walker per wavefront");
        for (int k = 0; k < P.K; ++k) {
            int s = gl * P.K + k;
            L.w[k * WAVE + gl] = s < P.n_off ? slot_weight(P, tb, L.x, s) : (T)0;
        }
    }
    int cnt = 0;
    if (P.has_jz) {
        const bool flags = P.flags;
        for (int e = gl; e < P.n_edges; e += G) {
            const int pk = P.edge_pack[e];
            unsigned char a = L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff];
            if (flags) L.anti[e] = a;
            cnt += a;
        }
    }
    return grp_sum_int<G>(cnt);
}

// lane sums (sequential) + group scan: returns this lane's inclusive sum
template <typename T, int G = WAVE>
__device__ __forceinline__ T lane_scan(const DevProblem &P, const WalkerLds<T> &L, int gl,
                                       T *lane_sum) {
    const int K = G == WAVE ? P.K : P.Kg[gidx<G>()];
    T acc = (T)0;
    if (std::is_same<T, float>::value && K % 32 == 0) {
        // FP32: 32 slots per chunk, every LDS read in flight before the
        // first add (the same sequential sum)
        for (int k0 = 0; k0 < K; k0 += 32) {
            T v[32];
#pragma unroll
            for (int k = 0; k < 32; ++k) v[k] = L.w[(k0 + k) * G + gl];
#pragma unroll
            for (int k = 0; k < 32; ++k) acc += v[k];
        }
    } else {
        for (int k = 0; k < K; ++k) acc += L.w[k * G + gl];
    }
    *lane_sum = acc;
    return grp_incl_scan<G>(acc);
}

// The last slot of weight > 0 at or before slot lim, G slots per round
// (slot t lives at LDS position (t % K) G + t / K); 0 when there is none.
template <typename T, int G = WAVE>
__device__ int last_positive_slot(const DevProblem &P, const WalkerLds<T> &L, int K, int lim,
                                  int lane) {
    const int gl = lane & (G - 1);
    if (lim > P.n_off - 1) lim = P.n_off - 1;
    for (int top = lim; top >= 0; top -= G) {
        const int t = top - gl;
        const bool pos = t >= 0 && L.w[wpos_g<G>(t, K)] > (T)0;
        const unsigned long long b = grp_ballot<G>(pos, lane);
        if (b) return top - (__ffsll((unsigned long long)b) - 1);
    }
    return 0;
}

// Inverse-CDF draw (sample_index, evolution.py:30-59) in blocked order over
// the group's G lanes; u in [0, R).  Lane j, the first whose inclusive sum
// exceeds u, holds the draw; inside its block of K slots the CDF is base +
// the group's tree scan of G consecutive slots at a time (one LDS read and
// one DPP scan per lane and chunk), base starting at lane j - 1's inclusive
// sum; the first slot whose CDF exceeds u is drawn.  A zero-weight slot that
// rounding selects gives way to the last positive slot before it, and where
// rounding leaves u past lane j's block the last positive slot of the block
// or before it is taken.  Mirrors orc_sample_blocked.
template <typename T, int G = WAVE>
__device__ int sample_slot(const DevProblem &P, const WalkerLds<T> &L, T incl, T lane_sum, T u,
                           int lane) {
    (void)lane_sum;
    const int gl = lane & (G - 1);
    const int K = G == WAVE ? P.K : P.Kg[gidx<G>()];
    const unsigned long long hit = grp_ballot<G>(incl > u, lane);
    const int j = hit ? __ffsll((unsigned long long)hit) - 1 : G - 1;
    T base = grp_bcast<G>(incl, j > 0 ? j - 1 : 0);
    if (j == 0) base = (T)0;
    for (int c0 = 0; c0 < K; c0 += G) {
        const int k = c0 + gl;
        const bool ok = k < K && j * K + k < P.n_off;
        const T v = ok ? L.w[k * G + j] : (T)0;
        const T cdf = base + grp_incl_scan<G>(v);
        const unsigned long long h = grp_ballot<G>(ok && cdf > u, lane);
        if (h) {
            const int g0 = __ffsll((unsigned long long)h) - 1;
            const int slot = j * K + c0 + g0;
            const unsigned long long pos = grp_ballot<G>(v > (T)0, lane);
            if ((pos >> g0) & 1ull) return slot;
            return last_positive_slot<T, G>(P, L, K, slot - 1, lane);
        }
        base = grp_bcast<G>(cdf, G - 1);
    }
    return last_positive_slot<T, G>(P, L, K, j * K + K - 1, lane);
}

// The Philox blocks of a walker's hops, G at a time: lane gl of the group
// holds the block of hop it0 + gl under the counters of gfmc_philox(it, wk,
// step, tag, key), so the ten rounds of G hops run once, on the vector unit,
// and hop it reads its block from group lane it % G (the same numbers as one
// gfmc_philox call per hop).
template <int G>
struct HopRng {
    gfmc_u4 v;
    uint32_t it0 = 0xffffffffu;

    __device__ __forceinline__ gfmc_u4 at(uint32_t it, uint32_t wk, uint32_t step,
                                          uint32_t tag, uint32_t k0, uint32_t k1, int gl) {
        const uint32_t base = it & ~(uint32_t)(G - 1);
        if (base != it0) {
            it0 = base;
            v = gfmc_philox(base + (uint32_t)gl, wk, step, tag, k0, k1);
        }
        const int src = (int)(it & (uint32_t)(G - 1));
        gfmc_u4 r;
        r.x = (uint32_t)grp_bcast<G>((int)v.x, src);
        r.y = (uint32_t)grp_bcast<G>((int)v.y, src);
        r.z = (uint32_t)grp_bcast<G>((int)v.z, src);
        r.w = (uint32_t)grp_bcast<G>((int)v.w, src);
        return r;
    }
};

// Data of move idx that the record path prefetches before the hold
// arithmetic: this lane's record, this lane's neighbour entry, the sites.
// The local-field form prefetches the two-word record and this lane's
// field update instead.
struct MovePrefetch {
    Rec rec;
    uint2 rf;
    int fnb;
    int nbr;
    int sites;
};

template <int FAST, int G>
__device__ __forceinline__ MovePrefetch prefetch_move(const DevProblem &P, int idx, int gl) {
    MovePrefetch m;
    if constexpr (FAST == 2) {
        m.rf = gl < P.R ? P.mv_rf_g[gidx<G>()][(size_t)idx * P.R + gl]
                        : make_uint2(0xffffffffu, 0xffffffffu);
        m.fnb = gl < P.FD ? P.mv_fnb[(size_t)idx * P.FD + gl] : P.N | (P.N << 16);
    } else {
        m.rec = gl < P.R ? load_rec(P.mv_rec, (size_t)idx * P.R + gl) : empty_rec();
    }
    m.nbr = (P.has_jz && P.uniform_jz && gl < P.RN) ? P.mv_nbr[(size_t)idx * P.RN + gl] : -1;
    m.sites = P.mv_sites[idx];
    return m;
}

// record path of apply_move: count the edges that turn (anti-)parallel,
// flip, re-evaluate the prefetched records
template <typename T, int FAST, int G = WAVE>
__device__ int apply_move_rec(const DevProblem &P, const Tables<T> &tb, WalkerLds<T> &L,
                              const MovePrefetch &m, int idx, int n_anti, int lane) {
    const int gl = lane & (G - 1);
    if (P.has_jz && P.uniform_jz) {
        bool plus = false, minus = false;
        if (m.nbr >= 0) {
            const int p = m.nbr & 0xffff, f = (m.nbr >> 16) & 0xffff;
            plus = L.x[f] == L.x[p];
            minus = !plus;
        }
        n_anti += __popcll(grp_ballot<G>(plus, lane)) - __popcll(grp_ballot<G>(minus, lane));
    }
    wave_sync();
    if (gl == 0) {
        const int a = m.sites & 0xffff, b = (m.sites >> 16) & 0xffff;
        L.x[a] = (signed char)-L.x[a];
        if (b != 0xffff) L.x[b] = (signed char)-L.x[b];
    }
    wave_sync();
    if constexpr (FAST == 2) {
        // the fields: the partners of site i, then those of site j (a partner
        // of both takes both updates, one pass after the other)
        const int fp = m.fnb & 0xffff, ff = (m.fnb >> 16) & 0xffff;
        if (gl < P.max_deg) L.f[fp] = (signed char)(L.f[fp] + 2 * L.x[ff]);
        wave_sync();
        if (gl >= P.max_deg && gl < P.FD) L.f[fp] = (signed char)(L.f[fp] + 2 * L.x[ff]);
        wave_sync();
        if (m.rf.x != 0xffffffffu) L.w[m.rf.x & 0xffffu] = eval_recf(tb, m.rf, L.x, L.f);
    } else {
        const int s = rec_f(m.rec, 0);
        if (s >= 0)
            L.w[wpos_g<G>(s, G == WAVE ? P.K : P.Kg[gidx<G>()])] = eval_rec(P, tb, m.rec, L.x);
    }
    if (FAST != 2 && P.has_jz && !P.uniform_jz) {
        for (int t = gl; t < P.RE; t += G) {
            const int e = P.mv_edges[(size_t)idx * P.RE + t];
            if (e >= 0) {
                const int pk = P.edge_pack[e];
                L.anti[e] = L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff];
            }
        }
    }
    wave_sync();
    return n_anti;
}

// flip the sites of move idx and re-evaluate what depends on them; returns
// the new anti-parallel count
template <typename T>
__device__ int apply_move(const DevProblem &P, const Tables<T> &tb, WalkerLds<T> &L, int idx,
                          int n_anti, int lane) {
    if (lane == 0) {
        if (idx < P.n_hx) {
            L.x[idx] = (signed char)-L.x[idx];
        } else {
            int e = idx - P.n_hx;
            int a = P.edge_i[e], b = P.edge_j[e];
            L.x[a] = (signed char)-L.x[a];
            L.x[b] = (signed char)-L.x[b];
        }
    }
    __syncthreads();
    if (P.full) {
        int cnt = full_eval<T, false>(P, tb, L, lane);
        __syncthreads();
        return cnt;
    }
    const int o0 = P.dep_off[idx], o1 = P.dep_off[idx + 1];
    for (int t = o0 + lane; t < o1; t += WAVE) {
        int s = P.dep_slots[t];
        L.w[wpos(s, P.K)] = slot_weight(P, tb, L.x, s);
    }
    if (P.has_jz) {
        const int e0 = P.edge_off[idx], e1 = P.edge_off[idx + 1];
        int delta = 0;
        for (int t = e0 + lane; t < e1; t += WAVE) {
            int e = P.dep_edges[t];
            unsigned char a = L.x[P.edge_i[e]] != L.x[P.edge_j[e]];
            delta += (int)a - (int)L.anti[e];
            L.anti[e] = a;
        }
        n_anti += wave_sum_int(delta);
    }
    __syncthreads();
    return n_anti;
}

// ---------------------------------------------------------------------------
// propagation kernel (evolve_walker_cont_time, evolution.py:559-721)
// ---------------------------------------------------------------------------
// Propagators (evolution.py, GFMC_METHOD_* of gfmc.h): continuous time
// (evolve_walker_cont_time, :559-721), n_bra elementary sub-steps of
// (lambda - H) (single_step, :322-424) and the same with the diagonal stays
// of a hold drawn as one geometric variate (multi_step, :427-556).

// One walker's branching step by a group of G lanes (G = 64: the whole
// wavefront; the generic path only runs with G = 64): spins at sp, log_w /
// e_loc / evals at lw_p / el_p / ev_p, LDS state at lds; wk is the walker
// index in the Philox counter (hop, wk, step, tag).  step_size is beta
// (continuous time) or the sub-step count n_bra (discrete methods, shift
// P.lam).  Groups of one wavefront run independently (no workgroup
// barriers on this path).
template <typename T, int FAST, int G, int METHOD>
__device__ __forceinline__ void propagate_walker(const DevProblem &P, const Tables<T> &tb,
                                                 unsigned char *lds, int lane, uint32_t wk,
                                                 uint32_t *sp, double *lw_p, double *el_p,
                                                 int32_t *ev_p, double step_size, double alw,
                                                 uint32_t key0, uint32_t key1, uint32_t step,
                                                 uint32_t tag, const CarryDev &cy) {
    const int gl = lane & (G - 1);
    // one walker per wavefront: its index is wave-uniform, so the Philox
    // rounds of every hop run on the scalar unit
    if constexpr (G == WAVE) wk = __builtin_amdgcn_readfirstlane(wk);
    WalkerLds<T> L = FAST ? carve_g<T, G>(P, lds) : carve<T>(P, lds);
    load_spins<G>(P, sp, L.x, gl);
    int n_anti = 0;
    int cur = 0;
    bool carried = false;
    if constexpr (FAST) {
        if (cy.on) {
            cur = cy.state[1];
            carried = cy.state[0] != 0 && cy.state[2] == G;
            if (carried) {
                // the parent's row holds the spins it ended its step with;
                // a walker whose spins differ (changed outside the step
                // loop) re-evaluates everything
                const uint32_t *cs = (const uint32_t *)((const T *)cy.w[cur] +
                                                        (int64_t)cy.parents[wk] * cy.row +
                                                        cy.s_off);
                bool same = true;
                for (int q = gl; q < P.words; q += G) same = same && cs[q] == sp[q];
                carried = grp_ballot<G>(!same, lane) == 0ull;
            }
        }
    }
    if (carried) {
        // the parent's weights and anti-parallel count from the last step
        // 16-byte pieces: rows and the LDS block are 16-byte aligned, n is a
        // multiple of G >= 16 values
        const int n16 = P.Kg[gidx<G>()] * G * (int)sizeof(T) / 16;
        const int64_t par = cy.parents[wk];
        const uint4 *src = (const uint4 *)((const T *)cy.w[cur] + par * cy.row);
        uint4 *dst = (uint4 *)L.w;
#pragma unroll 4
        for (int q = gl; q < n16; q += G) dst[q] = src[q];
        if constexpr (FAST == 2) {
            // the parent's local fields (those of the spins just loaded) and
            // the zero spin
            const uint4 *fs = (const uint4 *)((const T *)cy.w[cur] + par * cy.row + cy.f_off);
            for (int q = gl; q < P.f_pad / 16; q += G) ((uint4 *)L.f)[q] = fs[q];
            if (P.zsite && gl == 0) L.x[P.N] = 0;
        }
        n_anti = cy.anti[cur][par];
        wave_sync();
    } else {
        if constexpr (FAST == 2) fld_init<G>(P, L.x, L.f, gl);
        wave_sync();
        n_anti = full_eval<T, FAST, G>(P, tb, L, gl);
    }
    wave_sync();
    double diag = diag_value<G>(P, L.x, L.anti, n_anti, gl);
    double lw = *lw_p;
    double eL = 0.0;
    uint32_t it = 0;
    HopRng<G> rng;
    // flip the sites of slot idx, re-evaluate what depends on them, new diag
    auto move = [&](int idx, const MovePrefetch &m) {
        if constexpr (FAST)
            n_anti = apply_move_rec<T, FAST, G>(P, tb, L, m, idx, n_anti, lane);
        else
            n_anti = apply_move(P, tb, L, idx, n_anti, lane);
        if constexpr (FAST == 2) {
            // the local-field form implies uniform J_z (or none) and no h_z
            diag = P.has_jz ? (P.jz0 / 4) * (double)((P.n_edges - n_anti) - n_anti) : 0.0;
        } else {
            if (!(P.has_jz && P.uniform_jz) || P.has_hz)
                diag = diag_value<G>(P, L.x, L.anti, n_anti, gl);
            else
                diag = (P.jz0 / 4) * (double)((P.n_edges - n_anti) - n_anti);
        }
    };
    if constexpr (METHOD == GFMC_METHOD_CONT) {
        const double beta = step_size;
        double t = 0.0;
        while (t < beta) {
            T lane_sum;
            T incl = lane_scan<T, G>(P, L, gl, &lane_sum);
            T R = grp_bcast<G>(incl, G - 1);
            eL = diag - (double)R;
            gfmc_u4 r = rng.at(it, wk, step, tag, key0, key1, gl);
            double psi = gfmc_u53(r.z, r.w);
            int idx = -1;
            double tau_raw = INFINITY;
            MovePrefetch m;
            if (hop_rate(R)) {
                T u = uniform_for<T>(r) * R;
                T cap = below(R);
                if (u > cap) u = cap;
                idx = sample_slot<T, G>(P, L, incl, lane_sum, u, lane);
                if constexpr (FAST) m = prefetch_move<FAST, G>(P, idx, gl);  // in flight during the hold math
                if constexpr (FAST == 2) {
                    // FP32 kernels: -log(1 - psi) / R in FP32 (v_log_f32), the
                    // precision the rates carry
                    const float nl = -0.69314718055994531f *
                                     __builtin_amdgcn_logf((float)(1.0 - psi));
                    tau_raw = (double)(nl * __builtin_amdgcn_rcpf((float)R));
                } else {
                    tau_raw = gfmc_neg_log_unit(1.0 - psi) / (double)R;
                }
            }
            double rem = beta - t;
            double tau = tau_raw < rem ? tau_raw : rem;
            lw = lw - eL * tau;
            t = t + tau;
            ++it;
            if (tau_raw < rem && idx >= 0) move(idx, m);
        }
        lw = lw - alw * beta;
    } else {
        const int n_bra = (int)step_size;
        if constexpr (METHOD == GFMC_METHOD_SINGLE) {
            // sub-step: stay with weight w0 = lam - H_xx or hop to slot s
            // with weight -rate_s; log_w += log(b_x), b_x = w0 + R.  The draw
            // u in [0, b_x) is tested against w0 first, then u - w0 against
            // the blocked CDF of the slots.
            for (; it < (uint32_t)n_bra;) {
                T lane_sum;
                T incl = lane_scan<T, G>(P, L, gl, &lane_sum);
                T R = grp_bcast<G>(incl, G - 1);
                const double w0d = P.lam - diag;
                const T w0 = (T)w0d;
                const T W = w0 + R;
                gfmc_u4 r = rng.at(it, wk, step, tag, key0, key1, gl);
                T u = uniform_for<T>(r) * W;
                const T cap = below(W);
                if (u > cap) u = cap;
                int idx = -1;
                MovePrefetch m;
                // a total W <= 0 stays (no slot has weight; the walker dies
                // through log W), also when there are no off-diagonal slots
                if (W > (T)0 && !(u < w0)) {
                    idx = sample_slot<T, G>(P, L, incl, lane_sum, u - w0, lane);
                    if constexpr (FAST) m = prefetch_move<FAST, G>(P, idx, gl);
                }
                lw = lw + log(w0d + (double)R);
                ++it;
                if (idx >= 0) move(idx, m);
            }
        } else {
            // hold: k sub-steps at once, P(stay) = 1 - R / b_x per sub-step;
            // k = min(floor(log1p(-psi) / log1p(-R / b_x)) + 1, remaining + 1)
            int used = 0;
            while (used < n_bra) {
                T lane_sum;
                T incl = lane_scan<T, G>(P, L, gl, &lane_sum);
                T R = grp_bcast<G>(incl, G - 1);
                const double bx = (P.lam - diag) + (double)R;
                gfmc_u4 r = rng.at(it, wk, step, tag, key0, key1, gl);
                const double psi = gfmc_u53(r.z, r.w);
                int idx = -1;
                MovePrefetch m;
                if (R > (T)0) {
                    T u = uniform_for<T>(r) * R;
                    T cap = below(R);
                    if (u > cap) u = cap;
                    idx = sample_slot<T, G>(P, L, incl, lane_sum, u, lane);
                    if constexpr (FAST) m = prefetch_move<FAST, G>(P, idx, gl);
                }
                const int remaining = n_bra - used;
                const double p_hop = (double)R / bx;
                double k_raw = (double)remaining + 1.0;
                if (p_hop > 0.0) {
                    const double hold = floor(log1p(-psi) / log1p(-p_hop)) + 1.0;
                    k_raw = hold < k_raw ? hold : k_raw;
                }
                const int kr = (int)k_raw;
                const int k = kr < remaining ? kr : remaining;
                lw = lw + log(bx) * (double)k;
                used += k;
                ++it;
                if (kr <= remaining && idx >= 0) move(idx, m);
            }
        }
        // e_L of the final configuration: one more evaluation
        T lane_sum;
        T incl = lane_scan<T, G>(P, L, gl, &lane_sum);
        T R = grp_bcast<G>(incl, G - 1);
        eL = diag - (double)R;
        ++it;
        lw = lw - alw * (double)n_bra;
    }
    store_spins<G>(P, sp, L.x, gl,
                   FAST && cy.on ? (uint32_t *)((T *)cy.w[cur ^ 1] + (int64_t)wk * cy.row +
                                                cy.s_off)
                                 : nullptr);
    if constexpr (FAST) {
        if (cy.on) {
            wave_sync();
            const int n16 = P.Kg[gidx<G>()] * G * (int)sizeof(T) / 16;
            uint4 *dst = (uint4 *)((T *)cy.w[cur ^ 1] + (int64_t)wk * cy.row);
            const uint4 *src = (const uint4 *)L.w;
#pragma unroll 4
            for (int q = gl; q < n16; q += G) dst[q] = src[q];
            if constexpr (FAST == 2) {
                uint4 *fd = (uint4 *)((T *)cy.w[cur ^ 1] + (int64_t)wk * cy.row + cy.f_off);
                for (int q = gl; q < P.f_pad / 16; q += G) fd[q] = ((const uint4 *)L.f)[q];
            }
            if (gl == 0) cy.anti[cur ^ 1][wk] = n_anti;
        }
    }
    if (gl == 0) {
        *lw_p = lw;
        *el_p = eL;
        if (ev_p) *ev_p += (int32_t)it;
    }
}

// after a propagation that stored its rows: they are the latest, for G lanes
__global__ void k_carry_commit(int32_t *state, int G) {
    if (threadIdx.x == 0) {
        state[0] = 1;
        state[1] ^= 1;
        state[2] = G;
    }
}

// P.wpb wavefronts per workgroup (the record paths; 1 on the generic path),
// 64 / G walkers per wavefront.  With P.xcd, workgroup b of a grid of nb
// (nb8 = nb rounded down to a multiple of 8) runs block (b % 8) nb8 / 8 + b / 8
// when b < nb8: the dispatcher hands workgroups to the eight XCDs in turn,
// so each XCD then holds a contiguous range of walkers (and the siblings
// that read one parent's carried row share that XCD's L2).
template <typename T, int FAST, int G, int METHOD>
__global__ __launch_bounds__(4 * WAVE) void k_propagate(DevProblem P, Tables<T> tb, int64_t nw,
                                                        uint32_t *spins, double *log_w,
                                                        double *e_loc, double beta, double alw,
                                                        uint32_t key0, uint32_t key1,
                                                        uint32_t step, uint32_t tag,
                                                        int32_t *evals,
                                                        const uint32_t *step_base, CarryDev cy) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int grp = lane / G;
    int64_t b = blockIdx.x;
    if (P.xcd) {
        const int64_t nb8 = (int64_t)gridDim.x & ~(int64_t)7;
        if (b < nb8) b = (b & 7) * (nb8 >> 3) + (b >> 3);
    }
    const int64_t wk = (b * P.wpb + wave) * (WAVE / G) + grp;
    if (wk >= nw) return;
    const uint32_t st = step + (step_base ? *step_base : 0u);
    unsigned char *base =
        FAST ? lds + (size_t)(wave * (WAVE / G) + grp) * walker_lds_bytes<T, G>(P) : lds;
    propagate_walker<T, FAST, G, METHOD>(P, tb, base, lane, (uint32_t)wk,
                                         spins + wk * P.words,
                                 log_w + wk, e_loc + wk, evals ? evals + wk : nullptr, beta,
                                 alw, key0, key1, st, tag, cy);
}

// A batch of independent lattices in one launch: walker g belongs to lattice
// l with woff[l] <= g < woff[l + 1] and runs exactly as walker g - woff[l] of
// a single-lattice launch with the Philox key seed + l and shift alw[l];
// spins of every lattice use the common stride `stride` words.  With G < 64
// every lattice's walker count is a multiple of 64 / G, so the lattice is
// uniform over a wavefront.
template <typename T, int FAST, int G, int METHOD>
__global__ __launch_bounds__(WAVE) void k_propagate_batch(
    const DevProblem *Ps, const Tables<T> *Ts, int n_lat, const int64_t *woff, int64_t n_total,
    int stride, uint32_t *spins, double *log_w, double *e_loc, double beta, const double *alw,
    uint64_t seed, uint32_t step, uint32_t tag, int32_t *evals, const uint32_t *step_base) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t g0 = (int64_t)blockIdx.x * (WAVE / G);
    const int grp = threadIdx.x / G;
    const int64_t g = g0 + grp;
    if (g >= n_total) return;
    step += step_base ? *step_base : 0u;
    int lo = 0, hi = n_lat;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (woff[mid] <= g0) lo = mid;
        else hi = mid;
    }
    const uint64_t key = seed + (uint64_t)lo;
    const DevProblem &P = Ps[lo];
    unsigned char *base = G == WAVE ? lds : lds + (size_t)grp * walker_lds_bytes<T, G>(P);
    const CarryDev no_carry{};
    propagate_walker<T, FAST, G, METHOD>(P, Ts[lo], base, threadIdx.x, (uint32_t)(g - woff[lo]),
                                 spins + g * stride, log_w + g, e_loc + g,
                                 evals ? evals + g : nullptr, beta, alw[lo], (uint32_t)key,
                                 (uint32_t)(key >> 32), step, tag, no_carry);
}

// ---------------------------------------------------------------------------
// rates / log psi kernels
// ---------------------------------------------------------------------------
template <typename T, int FAST>
__global__ __launch_bounds__(WAVE) void k_rates(DevProblem P, Tables<T> tb, int64_t nw,
                                                const uint32_t *spins, double *rates,
                                                double *e_loc, double *log_diff) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x;
    const int64_t wk = blockIdx.x;
    if (wk >= nw) return;
    WalkerLds<T> L = carve<T>(P, lds);
    load_spins(P, spins + wk * P.words, L.x, lane);
    __syncthreads();
    if constexpr (FAST == 2) fld_init<WAVE>(P, L.x, L.f, lane);
    int n_anti = full_eval<T, FAST>(P, tb, L, lane);
    __syncthreads();
    double diag = diag_value(P, L.x, L.anti, n_anti, lane);
    T lane_sum;
    T incl = lane_scan(P, L, lane, &lane_sum);
    T R = lane_bcast(incl, WAVE - 1);
    const int64_t M = P.n_off + 1;
    if (rates) {
        if (lane == 0) rates[wk * M] = diag;
        for (int s = lane; s < P.n_off; s += WAVE)
            rates[wk * M + 1 + s] = -(double)L.w[wpos(s, P.K)];
    }
    if (log_diff) {
        if (lane == 0) log_diff[wk * M] = 0.0;
        for (int s = lane; s < P.n_off; s += WAVE)
            log_diff[wk * M + 1 + s] = P.guide == GFMC_GUIDE_JASTROW
                                           ? (double)slot_log_diff(P, tb, L.x, s)
                                           : 0.0;
    }
    if (e_loc && lane == 0) e_loc[wk] = diag - (double)R;
}

// log psi = 0.5 * sum_k sum_q J_kq x_k x_other + sum_k h_k x_k (sequential)
__global__ void k_log_psi(DevProblem P, Tables<double> tb, int64_t nw, const uint32_t *spins,
                          double *out) {
    int64_t wk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (wk >= nw) return;
    const uint32_t *sp = spins + wk * P.words;
    if (P.guide != GFMC_GUIDE_JASTROW) {
        out[wk] = 0.0;
        return;
    }
    const int D = P.max_deg;
    double pair = 0.0, field = 0.0;
    for (int k = 0; k < P.N; ++k) {
        double xk = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1.0 : -1.0;
        for (int q = 0; q < D; ++q) {
            int o = P.jas_site_other[k * D + q];
            double xo = ((sp[o >> 5] >> (o & 31)) & 1u) ? 1.0 : -1.0;
            pair += tb.jas_site_J[k * D + q] * xk * xo;
        }
    }
    for (int k = 0; k < P.N; ++k) {
        double xk = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1.0 : -1.0;
        field += tb.jas_h[k] * xk;
    }
    out[wk] = 0.5 * pair + field;
}

// ---------------------------------------------------------------------------
// RBM guide: log psi = sum_k a_k x_k + sum_h lncosh(theta_h),
// theta_h = b_h + sum_k W_hk x_k, W in BF16 staged in LDS once per workgroup
// of RBM_WAVES walker-wavefronts.  A slot's log-ratio is
//   -2 sum_{k in S} a_k x_k - sum_h Delta_h + sum_h log(p_h e^{2 Delta_h} + q_h)
// with Delta_h = -2 sum_{k in S} W_hk x_k, p = (1 + tanh theta)/2,
// q = (1 - tanh theta)/2 (cosh(theta + Delta)/cosh theta = e^{-Delta}
// (p e^{2 Delta} + q)); sum_h Delta_h comes from the column sums of W, and the
// logarithm is taken once per four hidden units of a running product.
// Every slot is re-evaluated after every move (theta changes everywhere).
// ---------------------------------------------------------------------------
#define RBM_WAVES 8

__device__ __forceinline__ float bf16_at(const uint32_t *W, int idx) {
    const uint32_t w = W[idx >> 1];
    return __uint_as_float((idx & 1) ? (w & 0xffff0000u) : (w << 16));
}

struct RbmLds {
    signed char *x;
    float *theta, *p, *q, *w;
    int *list;
    unsigned char *anti;
};

__device__ __forceinline__ RbmLds rbm_carve(const DevProblem &P, unsigned char *lds, int wave) {
    unsigned char *base = lds + (size_t)P.w_words * 4 + (size_t)wave * P.rbm_wave_bytes;
    RbmLds L;
    L.w = (float *)base;
    L.theta = L.w + P.K * WAVE;
    L.p = L.theta + P.H;
    L.q = L.p + P.H;
    L.list = (int *)(L.q + P.H);
    L.x = (signed char *)(L.list + P.n_off);
    L.anti = (unsigned char *)(L.x + P.n_pad);
    return L;
}

__device__ __forceinline__ void rbm_pq(float th, float *p, float *q) {
    const float e = __expf(-2.0f * fabsf(th));
    const float inv = 1.0f / (1.0f + e);
    const float big = inv, small = e * inv;
    *p = th >= 0.0f ? big : small;
    *q = th >= 0.0f ? small : big;
}

// theta from scratch (k ascending), lanes over hidden units
__device__ void rbm_theta_init(const DevProblem &P, const uint32_t *Wl, RbmLds &L, int lane) {
    for (int h = lane; h < P.H; h += WAVE) {
        float acc = P.rbm_b[h];
        const int row = h * (P.w_words * 2 / P.H);
        for (int k = 0; k < P.N; ++k) acc += bf16_at(Wl, row + k) * (float)L.x[k];
        L.theta[h] = acc;
        rbm_pq(acc, &L.p[h], &L.q[h]);
    }
}

// theta += Delta of flipping site i (and j >= 0) at the current (pre-flip) x
__device__ void rbm_theta_move(const DevProblem &P, const uint32_t *Wl, RbmLds &L, int i, int j,
                               int lane) {
    const int Np = P.w_words * 2 / P.H;
    const float xi = (float)L.x[i], xj = j >= 0 ? (float)L.x[j] : 0.0f;
    for (int h = lane; h < P.H; h += WAVE) {
        float d = -2.0f * xi * bf16_at(Wl, h * Np + i);
        if (j >= 0) d += -2.0f * xj * bf16_at(Wl, h * Np + j);
        const float th = L.theta[h] + d;
        L.theta[h] = th;
        rbm_pq(th, &L.p[h], &L.q[h]);
    }
}

// log psi(x') - log psi(x) of flipping i (and j >= 0)
__device__ float rbm_log_ratio(const DevProblem &P, const uint32_t *Wl, const RbmLds &L, int i,
                               int j) {
    const int Np = P.w_words * 2 / P.H;
    const float L2E = 1.4426950408889634f;
    const float xi = (float)L.x[i], xj = j >= 0 ? (float)L.x[j] : 0.0f;
    const float ci = -4.0f * L2E * xi, cj = -4.0f * L2E * xj;
    float acc = 0.0f;
    for (int h0 = 0; h0 < P.H; h0 += 4) {
        float prod = 1.0f;
        const int hn = h0 + 4 < P.H ? 4 : P.H - h0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u < hn) {
                const int h = h0 + u;
                float a = ci * bf16_at(Wl, h * Np + i);
                if (j >= 0) a += cj * bf16_at(Wl, h * Np + j);
                prod *= L.p[h] * __builtin_amdgcn_exp2f(a) + L.q[h];
            }
        }
        if (prod > 1e-30f && prod < 1e30f) {
            acc += __builtin_amdgcn_logf(prod);
        } else {
            for (int u = 0; u < hn; ++u) {
                const int h = h0 + u;
                float a = ci * bf16_at(Wl, h * Np + i);
                if (j >= 0) a += cj * bf16_at(Wl, h * Np + j);
                acc += __builtin_amdgcn_logf(L.p[h] * __builtin_amdgcn_exp2f(a) + L.q[h]);
            }
        }
    }
    float sum_delta = -2.0f * xi * P.rbm_cs[i];
    float afield = -2.0f * P.rbm_a[i] * xi;
    if (j >= 0) {
        sum_delta += -2.0f * xj * P.rbm_cs[j];
        afield += -2.0f * P.rbm_a[j] * xj;
    }
    return afield - sum_delta + 0.69314718055994531f * acc;
}

__device__ __forceinline__ void slot_sites(const DevProblem &P, int s, int *i, int *j) {
    if (s < P.n_hx) {
        *i = s;
        *j = -1;
    } else {
        const int pk = P.edge_pack[s - P.n_hx];
        *i = pk & 0xffff;
        *j = (pk >> 16) & 0xffff;
    }
}

// every slot weight (active slots compacted first), returns the anti count
__device__ int rbm_eval_all(const DevProblem &P, const Tables<float> &tb, const uint32_t *Wl,
                            RbmLds &L, int lane) {
    int base = 0;
    for (int c = 0; c < P.n_off; c += WAVE) {
        const int s = c + lane;
        float bare = 0.0f;
        if (s < P.n_off) {
            if (s < P.n_hx) {
                bare = tb.hx_half[s];
            } else {
                const int pk = P.edge_pack[s - P.n_hx];
                bare = (L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff]) ? tb.jpm[s - P.n_hx]
                                                                      : tb.jpmpm[s - P.n_hx];
            }
        }
        const bool act = s < P.n_off && bare != 0.0f;
        if (s < P.n_off && !act) L.w[wpos(s, P.K)] = -bare;
        const unsigned long long m = __ballot(act);
        const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        if (act) L.list[pos] = s;
        base += __popcll(m);
    }
    wave_sync();
    for (int t = lane; t < base; t += WAVE) {
        const int s = L.list[t];
        int i, j;
        slot_sites(P, s, &i, &j);
        const float bare = s < P.n_hx ? tb.hx_half[s]
                                      : ((L.x[i] != L.x[j]) ? tb.jpm[s - P.n_hx]
                                                            : tb.jpmpm[s - P.n_hx]);
        L.w[wpos(s, P.K)] = -(bare * __expf(rbm_log_ratio(P, Wl, L, i, j)));
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

__device__ void rbm_stage_W(const DevProblem &P, uint32_t *Wl) {
    for (int t = threadIdx.x; t < P.w_words; t += blockDim.x) Wl[t] = P.rbm_W[t];
    __syncthreads();
}

__device__ void rbm_walker_start(const DevProblem &P, const Tables<float> &tb,
                                 const uint32_t *Wl, RbmLds &L, const uint32_t *sp, int lane) {
    for (int k = lane; k < P.N; k += WAVE) L.x[k] = ((sp[k >> 5] >> (k & 31)) & 1u) ? 1 : -1;
    for (int s = P.n_off + lane; s < P.K * WAVE; s += WAVE) L.w[wpos(s, P.K)] = 0.0f;
    wave_sync();
    rbm_theta_init(P, Wl, L, lane);
    wave_sync();
}

__global__ __launch_bounds__(RBM_WAVES * WAVE) void k_propagate_rbm(
    DevProblem P, Tables<float> tb, int64_t nw, uint32_t *spins, double *log_w, double *e_loc,
    double beta, double alw, uint32_t key0, uint32_t key1, uint32_t step, uint32_t tag,
    int32_t *evals, const uint32_t *step_base) {
    step += step_base ? *step_base : 0u;
    extern __shared__ __align__(16) unsigned char lds[];
    uint32_t *Wl = (uint32_t *)lds;
    rbm_stage_W(P, Wl);
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    RbmLds L = rbm_carve(P, lds, wave);
    WalkerLds<float> WL;
    WL.w = L.w;
    WL.x = L.x;
    WL.anti = L.anti;
    for (int64_t wk = (int64_t)blockIdx.x * RBM_WAVES + wave; wk < nw;
         wk += (int64_t)gridDim.x * RBM_WAVES) {
        uint32_t *sp = spins + wk * P.words;
        rbm_walker_start(P, tb, Wl, L, sp, lane);
        int n_anti = rbm_eval_all(P, tb, Wl, L, lane);
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
                rbm_theta_move(P, Wl, L, i, j, lane);
                wave_sync();
                if (lane == 0) {
                    L.x[i] = (signed char)-L.x[i];
                    if (j >= 0) L.x[j] = (signed char)-L.x[j];
                }
                wave_sync();
                n_anti = rbm_eval_all(P, tb, Wl, L, lane);
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

__global__ __launch_bounds__(RBM_WAVES * WAVE) void k_rates_rbm(
    DevProblem P, Tables<float> tb, int64_t nw, const uint32_t *spins, double *rates,
    double *e_loc, double *log_diff, double *log_psi) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint32_t *Wl = (uint32_t *)lds;
    rbm_stage_W(P, Wl);
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    RbmLds L = rbm_carve(P, lds, wave);
    WalkerLds<float> WL;
    WL.w = L.w;
    WL.x = L.x;
    WL.anti = L.anti;
    const int64_t M = P.n_off + 1;
    for (int64_t wk = (int64_t)blockIdx.x * RBM_WAVES + wave; wk < nw;
         wk += (int64_t)gridDim.x * RBM_WAVES) {
        rbm_walker_start(P, tb, Wl, L, spins + wk * P.words, lane);
        if (log_psi) {
            float acc = 0.0f;
            for (int h = lane; h < P.H; h += WAVE) {
                const float a = fabsf(L.theta[h]);
                acc += a + log1pf(__expf(-2.0f * a)) - 0.69314718055994531f;
            }
            for (int k = lane; k < P.N; k += WAVE) acc += P.rbm_a[k] * (float)L.x[k];
            const float tot = lane_bcast(wave_incl_scan(acc, lane), WAVE - 1);
            if (lane == 0) log_psi[wk] = (double)tot;
        }
        if (rates || e_loc || log_diff) {
            const int n_anti = rbm_eval_all(P, tb, Wl, L, lane);
            const double diag = diag_value(P, L.x, L.anti, n_anti, lane);
            float lane_sum;
            const float incl = lane_scan(P, WL, lane, &lane_sum);
            const float R = lane_bcast(incl, WAVE - 1);
            if (rates) {
                if (lane == 0) rates[wk * M] = diag;
                for (int s = lane; s < P.n_off; s += WAVE)
                    rates[wk * M + 1 + s] = -(double)L.w[wpos(s, P.K)];
            }
            if (log_diff) {
                if (lane == 0) log_diff[wk * M] = 0.0;
                for (int s = lane; s < P.n_off; s += WAVE) {
                    int i, j;
                    slot_sites(P, s, &i, &j);
                    log_diff[wk * M + 1 + s] = (double)rbm_log_ratio(P, Wl, L, i, j);
                }
            }
            if (e_loc && lane == 0) e_loc[wk] = diag - (double)R;
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// initialisers, packing
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WAVE) void k_init_sz(int N, int words, int64_t nw, int n_up,
                                                  uint32_t key0, uint32_t key1,
                                                  uint32_t *spins) {
    extern __shared__ __align__(16) unsigned char lds[];
    signed char *x = (signed char *)lds;
    const int lane = threadIdx.x;
    const int64_t wk = blockIdx.x;
    if (wk >= nw) return;
    for (int k = lane; k < N; k += WAVE) x[k] = k < N - n_up ? -1 : 1;
    __syncthreads();
    if (lane == 0) {
        gfmc_u4 r = {0, 0, 0, 0};
        for (int i = N - 1; i >= 1; --i) {
            if (i == N - 1 || (i & 3) == 3)
                r = gfmc_philox((uint32_t)(i >> 2), (uint32_t)wk, 0u, GFMC_TAG_INIT, key0, key1);
            uint32_t rv = (i & 3) == 0 ? r.x : (i & 3) == 1 ? r.y : (i & 3) == 2 ? r.z : r.w;
            uint32_t j = (uint32_t)(((uint64_t)rv * (uint64_t)(i + 1)) >> 32);
            signed char tmp = x[i];
            x[i] = x[j];
            x[j] = tmp;
        }
    }
    __syncthreads();
    DevProblem P;
    P.N = N;
    P.words = words;
    store_spins(P, spins + wk * words, x, lane);
}

__global__ void k_init_random(int N, int words, int64_t nw, uint32_t key0, uint32_t key1,
                              uint32_t *spins) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nw * words) return;
    int64_t wk = g / words;
    int q = (int)(g % words);
    gfmc_u4 r = gfmc_philox((uint32_t)(q >> 2), (uint32_t)wk, 1u, GFMC_TAG_INIT, key0, key1);
    uint32_t v = (q & 3) == 0 ? r.x : (q & 3) == 1 ? r.y : (q & 3) == 2 ? r.z : r.w;
    int rem = N - 32 * q;
    if (rem < 32) v &= (1u << rem) - 1u;
    spins[g] = v;
}

__global__ void k_pack(int N, int words, int64_t nw, const double *config, uint32_t *spins) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nw * words) return;
    int64_t wk = g / words;
    int q = (int)(g % words);
    uint32_t v = 0;
    for (int b = 0; b < 32; ++b) {
        int k = q * 32 + b;
        if (k < N && config[wk * N + k] > 0) v |= 1u << b;
    }
    spins[g] = v;
}

__global__ void k_unpack(int N, int words, int64_t nw, const uint32_t *spins, double *config) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nw * N) return;
    int64_t wk = g / N;
    int k = (int)(g % N);
    config[g] = ((spins[wk * words + (k >> 5)] >> (k & 31)) & 1u) ? 1.0 : -1.0;
}

// ---------------------------------------------------------------------------
// reconfiguration (reconfiguration.py:9-62): six short kernels over blocks
// of RB = 1024 walkers (256 threads x 4 contiguous walkers).  Summation
// order: each thread adds its 4 walkers sequentially from 0; the 256 thread
// values of a block are combined by the wave tree inside each wave and
// sequentially over the 4 wave totals; block totals are combined
// sequentially in block order.  Scratch layout (doubles) after the n-entry
// CDF: bm[B] bs[B] bw[B] be[B] bml[B] blp[B] logG[1].
// ---------------------------------------------------------------------------
#define RC_T 256
#define RC_PER 4
#define RB (RC_T * RC_PER)

struct RcScratch {
    double *cdf, *bm, *bs, *bw, *be, *bml, *blp, *logG;
};

__device__ __forceinline__ RcScratch rc_scratch(double *work, int64_t n, int B) {
    RcScratch s;
    s.cdf = work;
    s.bm = work + n;
    s.bs = s.bm + B;
    s.bw = s.bs + B;
    s.be = s.bw + B;
    s.bml = s.be + B;
    s.blp = s.bml + B;
    s.logG = s.blp + B;
    return s;
}

// Segments: independent populations reconfigured in one launch (lattices of
// a batch).  nseg == 0: one population of n walkers, no tables.  Otherwise
// segment g holds walkers [w[g], w[g+1]) and blocks [b[g], b[g+1]) (each
// segment starts a new block), and its scratch of n_g + 7 B_g + 1 doubles at
// work + w[g] + 7 b[g] + g.  Every segment is reconfigured exactly as a
// population of its own.
struct RcSegs {
    int nseg;
    int64_t n;
    const int64_t *w;
    const int32_t *b;
};

struct RcLoc {
    int seg, lb, B;
    int64_t w0, n;
    double *work;
};

__device__ __forceinline__ RcLoc rc_locate(const RcSegs &S, double *work, int bid) {
    RcLoc r;
    if (S.nseg == 0) {
        r.seg = 0;
        r.w0 = 0;
        r.n = S.n;
        r.lb = bid;
        r.B = (int)((S.n + RB - 1) / RB);
        r.work = work;
        return r;
    }
    int lo = 0, hi = S.nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (S.b[mid] <= bid) lo = mid;
        else hi = mid;
    }
    r.seg = lo;
    r.w0 = S.w[lo];
    r.n = S.w[lo + 1] - r.w0;
    r.lb = bid - S.b[lo];
    r.B = S.b[lo + 1] - S.b[lo];
    r.work = work + r.w0 + 7 * (int64_t)S.b[lo] + lo;
    return r;
}

// inclusive scan of per-thread values in block order
__device__ double rc_block_incl(double v, double *sh4) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double incl = wave_incl_scan(v, lane);
    if (lane == 63) sh4[wv] = incl;
    __syncthreads();
    double P = 0.0;
    for (int q = 0; q < wv; ++q) P = P + sh4[q];
    __syncthreads();
    return P + incl;
}

__device__ double rc_block_max(double v, double *sh4) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        double o = __shfl_xor(v, d, WAVE);
        v = o > v ? o : v;
    }
    if (lane == 0) sh4[wv] = v;
    __syncthreads();
    double m = sh4[0];
    for (int q = 1; q < RC_T / WAVE; ++q) m = sh4[q] > m ? sh4[q] : m;
    __syncthreads();
    return m;
}

// exclusive running max over threads in block order (exact in any tree)
__device__ double rc_block_excl_max(double v, double *sh4) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double incl = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        double o = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl = o > incl ? o : incl;
    }
    double excl = __shfl_up(incl, 1, WAVE);
    if (lane == 0) excl = -INFINITY;
    if (lane == 63) sh4[wv] = incl;
    __syncthreads();
    double P = -INFINITY;
    for (int q = 0; q < wv; ++q) P = sh4[q] > P ? sh4[q] : P;
    __syncthreads();
    return P > excl ? P : excl;
}

__device__ __forceinline__ double rc_max_of(const double *a, int B) {
    double m = -INFINITY;
    for (int b = 0; b < B; ++b) m = a[b] > m ? a[b] : m;
    return m;
}

__device__ __forceinline__ double rc_m0(const RcScratch &s, int B) {
    const double m = rc_max_of(s.bm, B);
    return isfinite(m) ? m : 0.0;   // jax logsumexp replaces a non-finite max by 0
}

// 1: block maxima of log_w
__global__ __launch_bounds__(RC_T) void k_rc_max(RcSegs S, const double *lw, double *work) {
    __shared__ double sh4[RC_T / WAVE];
    const RcLoc g = rc_locate(S, work, blockIdx.x);
    RcScratch s = rc_scratch(g.work, g.n, g.B);
    lw += g.w0;
    const int64_t k0 = (int64_t)g.lb * RB + threadIdx.x * RC_PER;
    double m = -INFINITY;
#pragma unroll
    for (int c = 0; c < RC_PER; ++c)
        if (k0 + c < g.n) m = lw[k0 + c] > m ? lw[k0 + c] : m;
    m = rc_block_max(m, sh4);
    if (threadIdx.x == 0) s.bm[g.lb] = m;
}

// 2: block sums of exp(log_w - m0)
__global__ __launch_bounds__(RC_T) void k_rc_lse(RcSegs S, const double *lw, double *work) {
    __shared__ double sh4[RC_T / WAVE];
    const RcLoc g = rc_locate(S, work, blockIdx.x);
    RcScratch s = rc_scratch(g.work, g.n, g.B);
    lw += g.w0;
    const double m0 = rc_m0(s, g.B);
    const int64_t k0 = (int64_t)g.lb * RB + threadIdx.x * RC_PER;
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < RC_PER; ++c)
        if (k0 + c < g.n) acc += exp(lw[k0 + c] - m0);
    const double incl = rc_block_incl(acc, sh4);
    if (threadIdx.x == RC_T - 1) s.bs[g.lb] = incl;
}

__device__ __forceinline__ double rc_logG(const RcScratch &s, int B) {
    double S = 0.0;
    for (int b = 0; b < B; ++b) S = S + s.bs[b];
    return log(S) + rc_m0(s, B);
}

// 3: block sums of w and e w, last positive walker per block
__global__ __launch_bounds__(RC_T) void k_rc_weights(RcSegs S, const double *lw,
                                                    const double *el, double *work) {
    __shared__ double sh4[RC_T / WAVE];
    const RcLoc g = rc_locate(S, work, blockIdx.x);
    RcScratch s = rc_scratch(g.work, g.n, g.B);
    lw += g.w0;
    el += g.w0;
    const double logG = rc_logG(s, g.B);
    const int64_t k0 = (int64_t)g.lb * RB + threadIdx.x * RC_PER;
    double aw = 0.0, ae = 0.0, lp = -1.0;
#pragma unroll
    for (int c = 0; c < RC_PER; ++c) {
        if (k0 + c < g.n) {
            const double w = exp(lw[k0 + c] - logG);
            aw += w;
            ae += el[k0 + c] * w;
            if (w > 0) lp = (double)(k0 + c);
        }
    }
    const double iw = rc_block_incl(aw, sh4);
    const double ie = rc_block_incl(ae, sh4);
    lp = rc_block_max(lp, sh4);
    if (threadIdx.x == RC_T - 1) {
        s.bw[g.lb] = iw;
        s.be[g.lb] = ie;
    }
    if (threadIdx.x == 0) {
        s.blp[g.lb] = lp;
        if (g.lb == 0) s.logG[0] = logG;
    }
}

// 4: raw prefix sums of the normalised weights, block maxima of them
__global__ __launch_bounds__(RC_T) void k_rc_cdf(RcSegs S, const double *lw, double *work) {
    __shared__ double sh4[RC_T / WAVE];
    __shared__ double sh_incl[RC_T];
    const RcLoc g = rc_locate(S, work, blockIdx.x);
    RcScratch s = rc_scratch(g.work, g.n, g.B);
    lw += g.w0;
    const double logG = s.logG[0];
    double Pb = 0.0;
    for (int b = 0; b < g.lb; ++b) Pb = Pb + s.bw[b];
    const int64_t k0 = (int64_t)g.lb * RB + threadIdx.x * RC_PER;
    double w[RC_PER], aw = 0.0;
#pragma unroll
    for (int c = 0; c < RC_PER; ++c) {
        w[c] = (k0 + c < g.n) ? exp(lw[k0 + c] - logG) : 0.0;
        if (k0 + c < g.n) aw += w[c];
    }
    const double incl = rc_block_incl(aw, sh4);
    sh_incl[threadIdx.x] = incl;
    __syncthreads();
    const double Et = threadIdx.x == 0 ? 0.0 : sh_incl[threadIdx.x - 1];
    double cv = Pb + Et;
    double last = -INFINITY;
#pragma unroll
    for (int c = 0; c < RC_PER; ++c) {
        if (k0 + c < g.n) {
            cv = cv + w[c];
            s.cdf[k0 + c] = cv;
            last = cv;
        }
    }
    last = rc_block_max(last, sh4);
    if (threadIdx.x == 0) s.bml[g.lb] = last;
}

// 5: running max, so the CDF is non-decreasing despite rounding
__global__ __launch_bounds__(RC_T) void k_rc_runmax(RcSegs S, double *work) {
    __shared__ double sh4[RC_T / WAVE];
    const RcLoc g = rc_locate(S, work, blockIdx.x);
    RcScratch s = rc_scratch(g.work, g.n, g.B);
    double Mb = -INFINITY;
    for (int b = 0; b < g.lb; ++b) Mb = s.bml[b] > Mb ? s.bml[b] : Mb;
    const int64_t k0 = (int64_t)g.lb * RB + threadIdx.x * RC_PER;
    double v[RC_PER], last = -INFINITY;
#pragma unroll
    for (int c = 0; c < RC_PER; ++c) {
        v[c] = (k0 + c < g.n) ? s.cdf[k0 + c] : -INFINITY;
        if (k0 + c < g.n) last = v[c];
    }
    double prev = rc_block_excl_max(last, sh4);
    prev = prev > Mb ? prev : Mb;
#pragma unroll
    for (int c = 0; c < RC_PER; ++c)
        if (k0 + c < g.n && v[c] < prev) s.cdf[k0 + c] = prev;
}

// 6: parents[i] = min(first k with cdf[k] > (i + psi) / n, last positive),
// thread per walker; segment g draws psi with the key seed + g and writes
// global indices w[g] + k and out2[2 g .. 2 g + 1]
__global__ __launch_bounds__(RC_T) void k_rc_parents(RcSegs S, int64_t n_total,
                                                     const double *work_c, uint64_t seed,
                                                     uint32_t step, uint32_t tag, int use_psi,
                                                     double psi_in, int32_t *parents,
                                                     double *out2, const uint32_t *step_base) {
    const int64_t gi = (int64_t)blockIdx.x * RC_T + threadIdx.x;
    if (gi >= n_total) return;
    step += step_base ? *step_base : 0u;
    int seg = 0;
    int64_t w0 = 0, n = S.n;
    int B, b0 = 0;
    if (S.nseg == 0) {
        B = (int)((n + RB - 1) / RB);
    } else {
        int lo = 0, hi = S.nseg;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (S.w[mid] <= gi) lo = mid;
            else hi = mid;
        }
        seg = lo;
        w0 = S.w[lo];
        n = S.w[lo + 1] - w0;
        b0 = S.b[lo];
        B = S.b[lo + 1] - b0;
    }
    double *work = const_cast<double *>(work_c) + (S.nseg ? w0 + 7 * (int64_t)b0 + seg : 0);
    RcScratch s = rc_scratch(work, n, B);
    double psi;
    if (use_psi) {
        psi = psi_in;
    } else {
        const uint64_t key = seed + (uint64_t)seg;
        gfmc_u4 r = gfmc_philox(0u, 0u, step, tag, (uint32_t)key, (uint32_t)(key >> 32));
        psi = gfmc_u53(r.x, r.y);
    }
    const double lpd = rc_max_of(s.blp, B);
    const int64_t lastpos = lpd < 0 ? 0 : (int64_t)lpd;
    const int64_t i = gi - w0;
    const double u = ((double)i + psi) / (double)n;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (s.cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    parents[gi] = (int32_t)(w0 + (lo < lastpos ? lo : lastpos));
    if (i == 0) {
        double En = 0.0;
        for (int b = 0; b < B; ++b) En = En + s.be[b];
        out2[2 * seg] = s.logG[0] - log((double)n);
        out2[2 * seg + 1] = En;
    }
}

__global__ void k_gather(int words, int64_t nw, const uint32_t *spins_in, const double *e_in,
                         const int32_t *parents, uint32_t *spins_out, double *e_out,
                         double *lw_out) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nw * words) return;
    int64_t wk = g / words;
    int q = (int)(g % words);
    int64_t p = parents[wk];
    spins_out[g] = spins_in[p * words + q];
    if (q == 0) {
        if (e_out) e_out[wk] = e_in[p];
        if (lw_out) lw_out[wk] = 0.0;
    }
}

// wave-tree test hook: one wave scans 64 values
template <typename T>
__global__ __launch_bounds__(WAVE) void k_wave_scan(const double *in, double *out) {
    const int lane = threadIdx.x;
    T v = (T)in[lane];
    out[lane] = (double)wave_incl_scan(v, lane);
}

// sampler test hook: one wave, weights in LDS in the propagator's layout
template <typename T>
__global__ __launch_bounds__(WAVE) void k_sample(int n, const double *weights, int n_draws,
                                                 const double *uniforms, int32_t *out) {
    extern __shared__ __align__(16) unsigned char lds[];
    DevProblem P;
    P.n_off = n;
    P.K = n > 0 ? (n + WAVE - 1) / WAVE : 1;
    P.n_hx = 0;
    P.n_pad = 0;
    WalkerLds<T> L = carve<T>(P, lds);
    const int lane = threadIdx.x;
    for (int k = 0; k < P.K; ++k) {
        int s = lane * P.K + k;
        L.w[k * WAVE + lane] = s < n ? (T)weights[s] : (T)0;
    }
    __syncthreads();
    T lane_sum;
    T incl = lane_scan(P, L, lane, &lane_sum);
    T R = lane_bcast(incl, WAVE - 1);
    for (int d = 0; d < n_draws; ++d) {
        T u = (T)uniforms[d] * R;
        T cap = below(R);
        if (u > cap) u = cap;
        int idx = sample_slot(P, L, incl, lane_sum, u, lane);
        if (lane == 0) out[d] = idx;
    }
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
template <typename V>
static int upload(gfmc_problem *p, const V *src, size_t n, const V **dst) {
    if (n == 0 || src == nullptr) {
        *dst = nullptr;
        return 0;
    }
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, n * sizeof(V)));
    p->allocs.push_back(d);
    HIP_TRY(hipMemcpy(d, src, n * sizeof(V), hipMemcpyHostToDevice));
    *dst = (const V *)d;
    return 0;
}

static int upload_f32(gfmc_problem *p, const double *src, size_t n, const float **dst) {
    if (n == 0 || src == nullptr) {
        *dst = nullptr;
        return 0;
    }
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    return upload(p, tmp.data(), n, dst);
}

// Host-built tables of the record path (see DevProblem).
struct HostRecords {
    std::vector<uint32_t> slot_rec, slot_rec_lds, mv_rec;   // 8 words = 16 int16 per record
    std::vector<uint32_t> slot_rec_g[3];                    // LDS order of G = 64, 32, 16
    std::vector<int> mv_nbr, mv_sites, mv_edges, edge_pack;
    int R = 0, RN = 0, RE = 0;
    double J0 = 0.0, h0 = 0.0;
    // local-field form (DevProblem::fld)
    int fld = 0, FD = 0, zsite = 1;
    std::vector<uint32_t> slot_rf_g[3], mv_rf_g[3], fsite;   // 2 words per record, 4 per site
    std::vector<int> mv_fnb;
};

static int get16(const uint32_t *rec, int k) {
    const int v = (int)((rec[k >> 1] >> ((k & 1) * 16)) & 0xffffu);
    return (v ^ 0x8000) - 0x8000;
}

// The local-field form of the records: every record's partners (mask bits
// set) must be the Jastrow partners of its site(s), less the other flipped
// site where that one is a partner, and no site may list a partner twice.
// Builds the two-word records, the per-move field updates and the per-site
// partner lists (site N, a zero spin, as padding).
static void build_field_records(const gfmc_problem_desc *d, HostRecords &hr) {
    const int N = d->n_sites, n_off = d->n_hx + d->n_jpm, D = d->jas_max_deg;
    hr.fld = 0;
    if (d->guide != GFMC_GUIDE_JASTROW || D < 1 || D > 8 || N + 1 >= 32768) return;
    std::vector<std::vector<int>> nj(N);
    for (int k = 0; k < N; ++k) {
        for (int q = 0; q < D; ++q)
            if (d->jas_site_J[(size_t)k * D + q] != 0.0)
                nj[k].push_back(d->jas_site_other[(size_t)k * D + q]);
        std::vector<int> u = nj[k];
        std::sort(u.begin(), u.end());
        if (std::adjacent_find(u.begin(), u.end()) != u.end()) return;
    }
    std::vector<uint32_t> rf((size_t)n_off * 2);
    for (int s = 0; s < n_off; ++s) {
        const uint32_t *r = &hr.slot_rec[(size_t)s * 8];
        const int i = get16(r, 1), j = get16(r, 2), mask = get16(r, 3) & 0xffff;
        int flags = 0;
        for (int side = 0; side < (j >= 0 ? 2 : 1); ++side) {
            const int site = side ? j : i, other = side ? i : j;
            std::vector<int> got, want = nj[site];
            for (int q = 0; q < REC_D; ++q)
                if ((mask >> (side * REC_D + q)) & 1) got.push_back(get16(r, 4 + side * REC_D + q));
            if (other >= 0) {
                auto it = std::find(want.begin(), want.end(), other);
                if (it != want.end()) {
                    want.erase(it);
                    flags |= 1 << side;
                }
            }
            std::sort(got.begin(), got.end());
            std::sort(want.begin(), want.end());
            if (got != want) return;
        }
        rf[(size_t)s * 2] = (uint32_t)(s & 0xffff) | ((uint32_t)i << 16);
        rf[(size_t)s * 2 + 1] = (uint32_t)(j & 0xffff) | ((uint32_t)flags << 16);
    }
    for (int gi = 0; gi < 3; ++gi) {
        const int G = WAVE >> gi;
        const int K = (n_off + G - 1) / G;
        hr.slot_rf_g[gi].assign((size_t)K * G * 2, 0xffffffffu);
        for (int s = 0; s < n_off; ++s) {
            const size_t p = (size_t)((s % K) * G + s / K);
            hr.slot_rf_g[gi][p * 2] = rf[(size_t)s * 2];
            hr.slot_rf_g[gi][p * 2 + 1] = rf[(size_t)s * 2 + 1];
        }
    }
    // per move, in the first field the LDS position of the slot in each
    // group layout instead of the slot index
    for (int gi = 0; gi < 3; ++gi) {
        const int G = WAVE >> gi;
        const int K = (n_off + G - 1) / G;
        hr.mv_rf_g[gi].assign((size_t)n_off * hr.R * 2, 0xffffffffu);
        for (int s = 0; s < n_off; ++s)
            for (int t = d->dep_off[s]; t < d->dep_off[s + 1]; ++t) {
                const size_t o = ((size_t)s * hr.R + (t - d->dep_off[s])) * 2;
                const int slot = d->dep_slots[t];
                const uint32_t pos = (uint32_t)((slot % K) * G + slot / K);
                hr.mv_rf_g[gi][o] = (rf[(size_t)slot * 2] & 0xffff0000u) | pos;
                hr.mv_rf_g[gi][o + 1] = rf[(size_t)slot * 2 + 1];
            }
    }
    // a zero spin (site N) is needed only where a table entry pads: sites
    // with fewer than D partners, or site flips and pair moves in one problem
    bool regular = true;
    for (int k = 0; k < N; ++k) regular = regular && (int)nj[k].size() == D;
    hr.zsite = (!regular || (d->n_hx > 0 && d->n_jpm > 0)) ? 1 : 0;
    hr.FD = d->n_jpm > 0 ? 2 * D : D;
    const int Z = N | (N << 16);
    hr.mv_fnb.assign((size_t)n_off * hr.FD, Z);
    for (int s = 0; s < n_off; ++s) {
        const int a = hr.mv_sites[s] & 0xffff, b = (hr.mv_sites[s] >> 16) & 0xffff;
        for (int side = 0; side < 2; ++side) {
            const int f = side ? b : a;
            if (f == 0xffff) continue;
            for (size_t q = 0; q < nj[f].size(); ++q)
                hr.mv_fnb[(size_t)s * hr.FD + side * D + q] = nj[f][q] | (f << 16);
        }
    }
    hr.fsite.assign((size_t)N * 4, (uint32_t)Z);
    for (int k = 0; k < N; ++k)
        for (size_t q = 0; q < nj[k].size(); ++q) {
            uint32_t &w = hr.fsite[(size_t)k * 4 + q / 2];
            const int sh = (int)(q & 1) * 16;
            w = (w & ~(0xffffu << sh)) | ((uint32_t)nj[k][q] << sh);
        }
    hr.fld = 1;
}

static void put16(uint32_t *rec, int k, int v) {
    uint32_t &w = rec[k >> 1];
    const int sh = (k & 1) * 16;
    w = (w & ~(0xffffu << sh)) | ((uint32_t)(v & 0xffff) << sh);
}

// Decide whether the record path applies and build its tables; the packed
// edge list is always built.  Returns true for the record path.
static bool build_records(const gfmc_problem_desc *d, HostRecords &hr) {
    const int N = d->n_sites, E = d->n_edges, n_hx = d->n_hx, n_off = d->n_hx + d->n_jpm;
    hr.edge_pack.resize(E > 0 ? E : 1, 0);
    for (int e = 0; e < E; ++e) hr.edge_pack[e] = d->edge_i[e] | (d->edge_j[e] << 16);
    if (!d->dep_off || N >= 32768 || n_off >= 32768 || n_off == 0) return false;
    for (int e = 0; e < E && d->n_jpm; ++e)
        if (d->jpm[e] != d->jpm[0] || d->jpmpm[e] != d->jpmpm[0]) return false;
    for (int k = 0; k < N && n_hx; ++k)
        if (d->hx_half[k] != d->hx_half[0]) return false;
    const bool jas = d->guide == GFMC_GUIDE_JASTROW;
    const int D = jas ? d->jas_max_deg : 0;
    if (D > REC_D) return false;
    if (jas) {
        bool have = false;
        auto tied = [&](double v) {
            if (v == 0.0) return true;
            if (!have) {
                hr.J0 = v;
                have = true;
                return true;
            }
            return v == hr.J0;
        };
        for (int i = 0; i < N * D; ++i)
            if (!tied(d->jas_site_J[i])) return false;
        for (int i = 0; i < E * 2 * D; ++i)
            if (!tied(d->jas_edge_J[i])) return false;
        hr.h0 = d->jas_h[0];
        for (int k = 0; k < N; ++k)
            if (d->jas_h[k] != hr.h0) return false;
    }
    // one record per slot
    hr.slot_rec.assign((size_t)n_off * 8, 0u);
    for (int s = 0; s < n_off; ++s) {
        uint32_t *r = &hr.slot_rec[(size_t)s * 8];
        int i, j, mask = 0;
        if (s < n_hx) {
            i = s;
            j = -1;
        } else {
            i = d->edge_i[s - n_hx];
            j = d->edge_j[s - n_hx];
        }
        put16(r, 0, s);
        put16(r, 1, i);
        put16(r, 2, j);
        for (int q = 0; q < 2 * REC_D; ++q) put16(r, 4 + q, i);
        if (jas) {
            if (s < n_hx) {
                for (int q = 0; q < D; ++q) {
                    put16(r, 4 + q, d->jas_site_other[s * D + q]);
                    if (d->jas_site_J[s * D + q] != 0.0) mask |= 1 << q;
                }
            } else {
                const int e = s - n_hx;
                for (int side = 0; side < 2; ++side)
                    for (int q = 0; q < D; ++q) {
                        const size_t o = ((size_t)e * 2 + side) * D + q;
                        put16(r, 4 + side * REC_D + q, d->jas_edge_other[o]);
                        if (d->jas_edge_J[o] != 0.0) mask |= 1 << (side * REC_D + q);
                    }
            }
        }
        put16(r, 3, mask);
    }
    // the same records in LDS order: position (s % K) * G + s / K with
    // K = ceil(n_off / G), for walker groups of G = 64, 32, 16 lanes
    for (int gi = 0; gi < 3; ++gi) {
        const int G = WAVE >> gi;
        const int K = (n_off + G - 1) / G;
        hr.slot_rec_g[gi].assign((size_t)K * G * 8, 0xffffffffu);
        for (int s = 0; s < n_off; ++s)
            memcpy(&hr.slot_rec_g[gi][(size_t)((s % K) * G + s / K) * 8],
                   &hr.slot_rec[(size_t)s * 8], 8 * sizeof(uint32_t));
    }
    hr.slot_rec_lds = hr.slot_rec_g[0];
    // per-move record blocks
    for (int s = 0; s < n_off; ++s) hr.R = std::max(hr.R, d->dep_off[s + 1] - d->dep_off[s]);
    if (hr.R > WAVE) return false;
    hr.mv_rec.assign((size_t)n_off * hr.R * 8, 0xffffffffu);
    for (int s = 0; s < n_off; ++s)
        for (int t = d->dep_off[s]; t < d->dep_off[s + 1]; ++t)
            memcpy(&hr.mv_rec[((size_t)s * hr.R + (t - d->dep_off[s])) * 8],
                   &hr.slot_rec[(size_t)d->dep_slots[t] * 8], 8 * sizeof(uint32_t));
    // flipped sites and the neighbours whose bond state flips
    std::vector<std::vector<int>> adj(N);
    for (int e = 0; e < E; ++e) {
        adj[d->edge_i[e]].push_back(d->edge_j[e]);
        adj[d->edge_j[e]].push_back(d->edge_i[e]);
    }
    hr.mv_sites.resize(n_off);
    std::vector<std::vector<int>> nb(n_off);
    for (int s = 0; s < n_off; ++s) {
        int a, b;
        if (s < n_hx) {
            a = s;
            b = -1;
        } else {
            a = d->edge_i[s - n_hx];
            b = d->edge_j[s - n_hx];
        }
        hr.mv_sites[s] = a | ((b < 0 ? 0xffff : b) << 16);
        for (int f : {a, b}) {
            if (f < 0) continue;
            for (int p : adj[f])
                if (p != a && p != b) nb[s].push_back(p | (f << 16));
        }
        hr.RN = std::max(hr.RN, (int)nb[s].size());
    }
    if (d->has_jz && d->uniform_jz && hr.RN > WAVE) return false;
    hr.RN = std::max(hr.RN, 1);
    hr.mv_nbr.assign((size_t)n_off * hr.RN, -1);
    for (int s = 0; s < n_off; ++s)
        for (size_t t = 0; t < nb[s].size(); ++t) hr.mv_nbr[(size_t)s * hr.RN + t] = nb[s][t];
    // edges whose state a move may change (non-uniform J_z keeps flags)
    if (d->has_jz) {
        for (int s = 0; s < n_off; ++s)
            hr.RE = std::max(hr.RE, d->edge_off[s + 1] - d->edge_off[s]);
        hr.RE = std::max(hr.RE, 1);
        hr.mv_edges.assign((size_t)n_off * hr.RE, -1);
        for (int s = 0; s < n_off; ++s)
            for (int t = d->edge_off[s]; t < d->edge_off[s + 1]; ++t)
                hr.mv_edges[(size_t)s * hr.RE + (t - d->edge_off[s])] = d->dep_edges[t];
    } else {
        hr.mv_edges.assign(1, -1);
    }
    build_field_records(d, hr);
    return true;
}

template <typename T>
static void launch_rates(const gfmc_problem *p, const Tables<T> &tb, int64_t nw,
                         const uint32_t *spins, double *rates, double *e_loc, double *log_diff,
                         hipStream_t st) {
    if constexpr (std::is_same<T, float>::value) {
        if (p->dp.fld) {
            hipLaunchKernelGGL((k_rates<T, 2>), dim3((unsigned)nw), dim3(WAVE),
                               (size_t)p->lds_bytes, st, p->dp, tb, nw, spins, rates, e_loc,
                               log_diff);
            return;
        }
    }
    if (p->dp.fast)
        hipLaunchKernelGGL((k_rates<T, true>), dim3((unsigned)nw), dim3(WAVE),
                           (size_t)p->lds_bytes, st, p->dp, tb, nw, spins, rates, e_loc, log_diff);
    else
        hipLaunchKernelGGL((k_rates<T, false>), dim3((unsigned)nw), dim3(WAVE),
                           (size_t)p->lds_bytes, st, p->dp, tb, nw, spins, rates, e_loc, log_diff);
}

// grid: P.wpb wavefronts of 64 / G walkers per workgroup; lds: per workgroup
template <typename T, int FAST, int G>
static void launch_prop_g(const gfmc_problem *p, const Tables<T> &tb, int64_t nw,
                          uint32_t *spins, double *log_w, double *e_loc, double beta,
                          double alw, uint32_t k0, uint32_t k1, uint32_t step, uint32_t tag,
                          int32_t *evals, hipStream_t st) {
    CarryDev cy{};
    if (FAST && g_carry && g_carry->on && !p->dp.flags) cy = *g_carry;
    // at most 64 KiB of LDS per workgroup
    const int64_t wbytes = walker_lds_bytes<T, G>(p->dp);
    int wpb = FAST ? p->dp.wpb : 1;
    while (wpb > 1 && (int64_t)wpb * (WAVE / G) * wbytes > 64 * 1024) wpb >>= 1;
    const int64_t per_block = (int64_t)wpb * (WAVE / G);
    const unsigned grid = (unsigned)((nw + per_block - 1) / per_block);
    const size_t lds = FAST ? (size_t)(per_block * wbytes) : (size_t)p->lds_bytes;
    const dim3 block(wpb * WAVE);
    DevProblem dp = p->dp;   // the kernel indexes walkers with the launched shape
    dp.wpb = wpb;
    if (p->dp.method == GFMC_METHOD_SINGLE)
        hipLaunchKernelGGL((k_propagate<T, FAST, G, GFMC_METHOD_SINGLE>), dim3(grid), block,
                           lds, st, dp, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step,
                           tag, evals, g_step_base, cy);
    else if (p->dp.method == GFMC_METHOD_MULTI)
        hipLaunchKernelGGL((k_propagate<T, FAST, G, GFMC_METHOD_MULTI>), dim3(grid), block,
                           lds, st, dp, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step,
                           tag, evals, g_step_base, cy);
    else
        hipLaunchKernelGGL((k_propagate<T, FAST, G, GFMC_METHOD_CONT>), dim3(grid), block,
                           lds, st, dp, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step,
                           tag, evals, g_step_base, cy);
    if (cy.on) hipLaunchKernelGGL(k_carry_commit, dim3(1), dim3(WAVE), 0, st, cy.state, G);
}

// the record path at the problem's lanes per walker (FAST 1: records, 2:
// their local-field form)
template <typename T, int FAST>
static void launch_prop_rec(const gfmc_problem *p, const Tables<T> &tb, int64_t nw,
                            uint32_t *spins, double *log_w, double *e_loc, double beta,
                            double alw, uint32_t k0, uint32_t k1, uint32_t step, uint32_t tag,
                            int32_t *evals, hipStream_t st) {
    if (p->lanes == 32)
        launch_prop_g<T, FAST, 32>(p, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag,
                                   evals, st);
    else if (p->lanes == 16)
        launch_prop_g<T, FAST, 16>(p, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag,
                                   evals, st);
    else
        launch_prop_g<T, FAST, 64>(p, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag,
                                   evals, st);
}

template <typename T>
static void launch_propagate(const gfmc_problem *p, const Tables<T> &tb, int64_t nw,
                             uint32_t *spins, double *log_w, double *e_loc, double beta,
                             double alw, uint32_t k0, uint32_t k1, uint32_t step, uint32_t tag,
                             int32_t *evals, hipStream_t st) {
    if (!p->dp.fast) {
        launch_prop_g<T, 0, 64>(p, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag,
                                evals, st);
        return;
    }
    if constexpr (std::is_same<T, float>::value) {
        if (p->dp.fld) {
            launch_prop_rec<T, 2>(p, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag,
                                  evals, st);
            return;
        }
    }
    launch_prop_rec<T, 1>(p, tb, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag, evals,
                          st);
}

#include "gfmc_rbm.inl"

// the gfmc_rbm.inl kernels for M = Hp / 64 hidden-unit chunks per lane
template <int M>
static int rbh_set_lds_m(int bytes) {
    HIP_TRY(hipFuncSetAttribute((const void *)k_propagate_rbh<M>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    HIP_TRY(hipFuncSetAttribute((const void *)k_rates_rbh<M>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    return 0;
}

static int rbh_set_lds(int M, int bytes) {
    switch (M) {
    case 1: return rbh_set_lds_m<1>(bytes);
    case 2: return rbh_set_lds_m<2>(bytes);
    case 3: return rbh_set_lds_m<3>(bytes);
    case 4: return rbh_set_lds_m<4>(bytes);
    case 5: return rbh_set_lds_m<5>(bytes);
    case 6: return rbh_set_lds_m<6>(bytes);
    case 7: return rbh_set_lds_m<7>(bytes);
    case 8: return rbh_set_lds_m<8>(bytes);
    }
    return fail(GFMC_EUNSUP, "RBM: hidden layer too wide for the lanes-over-units kernels");
}

// persistent grid of the gfmc_rbm.inl kernels: as many workgroups as fit
static unsigned rbh_grid(const gfmc_problem *p, int64_t nw) {
    const int64_t need = (nw + RBH_WAVES - 1) / RBH_WAVES;
    const int64_t per_cu = (160 * 1024) / (p->rbh_lds_bytes > 0 ? p->rbh_lds_bytes : 1);
    const int64_t cap = (int64_t)p->n_cu * (per_cu > 0 ? per_cu : 1);
    return (unsigned)(need < cap ? need : cap);
}

template <int M>
static void rbh_launch_m(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                         double *e_loc, double beta, double alw, uint32_t k0, uint32_t k1,
                         uint32_t step, uint32_t tag, int32_t *evals, hipStream_t st) {
    hipLaunchKernelGGL(k_rbh_queue_reset, dim3(1), dim3(WAVE), 0, st, p->dp.rbm_queue);
    hipLaunchKernelGGL(k_propagate_rbh<M>, dim3(rbh_grid(p, nw)), dim3(RBH_WAVES * WAVE),
                       (size_t)p->rbh_lds_bytes, st, p->dp, p->t32, nw, spins, log_w, e_loc,
                       beta, alw, k0, k1, step, tag, evals, g_step_base);
}

template <int M>
static void rbh_rates_m(const gfmc_problem *p, int64_t nw, const uint32_t *spins, double *rates,
                        double *e_loc, double *log_diff, double *log_psi, hipStream_t st) {
    hipLaunchKernelGGL(k_rates_rbh<M>, dim3(rbh_grid(p, nw)), dim3(RBH_WAVES * WAVE),
                       (size_t)p->rbh_lds_bytes, st, p->dp, p->t32, nw, spins, rates, e_loc,
                       log_diff, log_psi);
}

// the tied-RBM kernel (gfmc_rbm_set_tied) for CP pair and CS single chunks
static void rbt_build_tables(const gfmc_problem *p, hipStream_t st) {
    const DevProblem &P = p->dp;
    const int64_t total = (int64_t)(2 * P.rbt_ncls + 1) * P.rbt_F * P.rbt_Q + P.rbt_Q;
    const unsigned g = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_rbt_tables, dim3(g), dim3(256), 0, st, P, P.rbt_raw);
    hipLaunchKernelGGL(k_rbt_pair, dim3(g), dim3(256), 0, st, total, P.rbt_D,
                       (const float *)P.rbt_raw, (float2 *)P.rbt_tab);
}

// the (CP, CS, PK) shapes the tied kernel is built for
#define RBT_SWITCH(CP, CS, PK, FN, ...)                                \
    [&]() -> int {                                                     \
        switch ((CP) * 8 + (CS) * 2 + ((PK) ? 1 : 0)) {                \
        case 2: return FN<0, 1, false>(__VA_ARGS__);                   \
        case 3: return FN<0, 1, true>(__VA_ARGS__);                    \
        case 4: return FN<0, 2, false>(__VA_ARGS__);                   \
        case 8: return FN<1, 0, false>(__VA_ARGS__);                   \
        case 10: return FN<1, 1, false>(__VA_ARGS__);                  \
        case 11: return FN<1, 1, true>(__VA_ARGS__);                   \
        case 12: return FN<1, 2, false>(__VA_ARGS__);                  \
        case 16: return FN<2, 0, false>(__VA_ARGS__);                  \
        case 18: return FN<2, 1, false>(__VA_ARGS__);                  \
        case 19: return FN<2, 1, true>(__VA_ARGS__);                   \
        case 20: return FN<2, 2, false>(__VA_ARGS__);                  \
        case 24: return FN<3, 0, false>(__VA_ARGS__);                  \
        case 26: return FN<3, 1, false>(__VA_ARGS__);                  \
        case 27: return FN<3, 1, true>(__VA_ARGS__);                   \
        case 28: return FN<3, 2, false>(__VA_ARGS__);                  \
        case 32: return FN<4, 0, false>(__VA_ARGS__);                  \
        }                                                              \
        return fail(GFMC_EUNSUP, "RBM: no tied kernel for this chunk shape"); \
    }()

static bool rbt_shape_ok(int cp, int cs) {
    return cp >= 0 && cs >= 0 && cp <= 4 && cs <= 2 && cp + cs >= 1 && 2 * cp + cs <= 8;
}

template <int CP, int CS, bool PK>
static int rbt_set_lds_m(int bytes) {
    HIP_TRY(hipFuncSetAttribute((const void *)k_propagate_rbt<CP, CS, PK>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    return 0;
}

// persistent grid: as many workgroups of rbt_waves walkers as fit
static unsigned rbt_grid(const gfmc_problem *p, int64_t nw) {
    const int64_t need = (nw + p->rbt_waves - 1) / p->rbt_waves;
    const int64_t per_cu = (160 * 1024) / (p->rbt_lds_bytes > 0 ? p->rbt_lds_bytes : 1);
    const int64_t cap = (int64_t)p->n_cu * (per_cu > 0 ? per_cu : 1);
    return (unsigned)(need < cap ? need : cap);
}

template <int CP, int CS, bool PK>
static int rbt_launch_m(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                        double *e_loc, double beta, double alw, uint32_t k0, uint32_t k1,
                        uint32_t step, uint32_t tag, int32_t *evals, hipStream_t st) {
    CarryDev cy{};
    if (g_carry && g_carry->on) cy = *g_carry;   // on only where rbt_carries
    hipLaunchKernelGGL(k_rbh_queue_reset, dim3(1), dim3(WAVE), 0, st, p->dp.rbm_queue);
    hipLaunchKernelGGL((k_propagate_rbt<CP, CS, PK>), dim3(rbt_grid(p, nw)),
                       dim3(p->rbt_waves * WAVE), (size_t)p->rbt_lds_bytes, st, p->dp, p->t32,
                       nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag, evals,
                       g_step_base, cy);
    if (cy.on)
        hipLaunchKernelGGL(k_rbt_carry_commit, dim3(1), dim3(WAVE), 0, st, cy.state,
                           p->dp.rbt_gen);
    return 0;
}

#define RBH_SWITCH(M, FN, ...)                      \
    do {                                            \
        switch (M) {                                \
        case 1: FN<1>(__VA_ARGS__); break;          \
        case 2: FN<2>(__VA_ARGS__); break;          \
        case 3: FN<3>(__VA_ARGS__); break;          \
        case 4: FN<4>(__VA_ARGS__); break;          \
        case 5: FN<5>(__VA_ARGS__); break;          \
        case 6: FN<6>(__VA_ARGS__); break;          \
        case 7: FN<7>(__VA_ARGS__); break;          \
        default: FN<8>(__VA_ARGS__); break;         \
        }                                           \
    } while (0)

// persistent RBM grid: as many workgroups as fit (one per CU at full LDS)
static unsigned rbm_grid(const gfmc_problem *p, int64_t nw) {
    const int64_t need = (nw + RBM_WAVES - 1) / RBM_WAVES;
    const int64_t per_cu = (160 * 1024) / (p->lds_bytes > 0 ? p->lds_bytes : 1);
    const int64_t cap = (int64_t)p->n_cu * (per_cu > 0 ? per_cu : 1);
    return (unsigned)(need < cap ? need : cap);
}

static void launch_reconfigure(const RcSegs &S, int64_t n_total, unsigned blocks,
                               const double *log_w, const double *e_loc, uint64_t seed,
                               uint32_t step, uint32_t tag, int use_psi, double psi,
                               int32_t *parents, double *out2, double *work, hipStream_t st) {
    hipLaunchKernelGGL(k_rc_max, dim3(blocks), dim3(RC_T), 0, st, S, log_w, work);
    hipLaunchKernelGGL(k_rc_lse, dim3(blocks), dim3(RC_T), 0, st, S, log_w, work);
    hipLaunchKernelGGL(k_rc_weights, dim3(blocks), dim3(RC_T), 0, st, S, log_w, e_loc, work);
    hipLaunchKernelGGL(k_rc_cdf, dim3(blocks), dim3(RC_T), 0, st, S, log_w, work);
    hipLaunchKernelGGL(k_rc_runmax, dim3(blocks), dim3(RC_T), 0, st, S, work);
    hipLaunchKernelGGL(k_rc_parents, dim3((unsigned)((n_total + RC_T - 1) / RC_T)), dim3(RC_T),
                       0, st, S, n_total, (const double *)work, seed, step, tag, use_psi, psi,
                       parents, out2, g_step_base);
}

// the walker groups of one wavefront fit 64 KiB of LDS
static bool lanes_lds_fits(const gfmc_problem *p, int lanes) {
    int64_t b = 0;
    if (lanes == 32)
        b = 2 * (int64_t)(p->precision == GFMC_FP64 ? walker_lds_bytes<double, 32>(p->dp)
                                                     : walker_lds_bytes<float, 32>(p->dp));
    else if (lanes == 16)
        b = 4 * (int64_t)(p->precision == GFMC_FP64 ? walker_lds_bytes<double, 16>(p->dp)
                                                     : walker_lds_bytes<float, 16>(p->dp));
    else
        b = p->lds_bytes;
    return b <= 64 * 1024;
}

// MLP guide (gfmc_mlp.inl)
static int mlp_rates(const gfmc_problem *p, int64_t nw, const uint32_t *spins, double *rates,
                     double *e_loc, double *log_diff, double *log_psi, hipStream_t st);
static int mlp_propagate(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                         double *e_loc, double beta, double alw, uint32_t k0, uint32_t k1,
                         uint32_t step, uint32_t tag, int32_t *evals, hipStream_t st);
#define MLP_NT 2048
#define MLP_SW 4

// BF16 bits of v (round to nearest even), magnitudes below 2^-24 flushed to 0
static uint16_t mlp_bf16_bits(float v) {
    uint32_t bits;
    memcpy(&bits, &v, 4);
    uint32_t r = (bits + 0x7fffu + ((bits >> 16) & 1u)) >> 16;
    const uint32_t back = r << 16;
    float f;
    memcpy(&f, &back, 4);
    if (std::fabs(f) < 5.9604644775390625e-08f) r = 0;
    return (uint16_t)r;
}

// v on the MLP's layer-1 grid (W1, b1): BF16 (round to nearest even) with
// magnitudes below 2^-9 flushed to 0 and above 64 clamped to +-64, so that
// v x 2^16 is an integer of magnitude at most 2^22
static float mlp_l1_value(float v) {
    uint32_t bits;
    memcpy(&bits, &v, 4);
    const uint32_t back = ((bits + 0x7fffu + ((bits >> 16) & 1u)) >> 16) << 16;
    float f;
    memcpy(&f, &back, 4);
    if (std::fabs(f) < 0.001953125f) return 0.0f;
    if (std::fabs(f) > 64.0f) return std::copysign(64.0f, f);
    return f;
}

// diagonal forward walking (gfmc_diag.inl)
struct gfmc_diag;
static int diag_check(const char *fn, const gfmc_diag *d, const gfmc_problem *p, void *ring,
                      double *dwork);
static int64_t diag_row_doubles(const gfmc_diag *d);
static int diag_record_impl(const gfmc_diag *d, int64_t nw, const uint32_t *spins,
                            const int32_t *parents, void *ring, double *row, double *work,
                            hipStream_t st);

static void free_problem(gfmc_problem *p) {
    free_mlp_work(p->mlp);
    free_mlp_train(p->mlpt);
    for (void *a : p->allocs) (void)hipFree(a);
    delete p;
}

extern "C" {

const char *gfmc_version(void) { return "pygfmc_amd gfmc_hip 0.1 (gfx950)"; }
const char *gfmc_last_error(void) { return g_err.c_str(); }

int gfmc_problem_create(const gfmc_problem_desc *d, int32_t precision, gfmc_problem **out) {
    if (!d || !out) return fail(GFMC_EINVAL, "gfmc_problem_create: NULL argument");
    if (precision != GFMC_FP64 && precision != GFMC_FP32)
        return fail(GFMC_EINVAL, "gfmc_problem_create: precision must be GFMC_FP64 or GFMC_FP32");
    if (d->n_sites <= 0) return fail(GFMC_EINVAL, "gfmc_problem_create: n_sites must be > 0");
    if (d->n_sites > 65535)
        return fail(GFMC_EUNSUP, "gfmc_problem_create: at most 65535 sites per problem");
    if (d->n_edges < 0 || (d->n_edges > 0 && (!d->edge_i || !d->edge_j)))
        return fail(GFMC_EINVAL, "gfmc_problem_create: bad edge list");
    if (!(d->n_hx == 0 || d->n_hx == d->n_sites))
        return fail(GFMC_EINVAL, "gfmc_problem_create: n_hx must be 0 or n_sites");
    if (!(d->n_jpm == 0 || d->n_jpm == d->n_edges))
        return fail(GFMC_EINVAL, "gfmc_problem_create: n_jpm must be 0 or n_edges");
    if (d->guide == GFMC_GUIDE_RBM) {
        if (precision != GFMC_FP32)
            return fail(GFMC_EINVAL, "gfmc_problem_create: the RBM guide runs in FP32 "
                                     "(BF16 weights); pass GFMC_FP32");
        if (d->rbm_hidden <= 0 || !d->rbm_W || !d->rbm_a || !d->rbm_b)
            return fail(GFMC_EINVAL, "gfmc_problem_create: incomplete RBM parameters");
    }
    if (d->guide == GFMC_GUIDE_MLP) {
        if (precision != GFMC_FP32)
            return fail(GFMC_EINVAL, "gfmc_problem_create: the MLP guide runs in FP32 "
                                     "(BF16 layers); pass GFMC_FP32");
        if (d->mlp_h1 <= 0 || d->mlp_h1 % 64 != 0 || d->mlp_h1 > 4096 ||
            !(d->mlp_h2 == 32 || d->mlp_h2 == 64 || d->mlp_h2 == 128 || d->mlp_h2 == 256))
            return fail(GFMC_EINVAL, "gfmc_problem_create: MLP hidden sizes must be a multiple "
                                     "of 64 (layer 1, at most 4096) and 32, 64, 128 or 256 "
                                     "(layer 2)");
        if (!d->mlp_W1 || !d->mlp_b1 || !d->mlp_W2 || !d->mlp_b2 || !d->mlp_w3 || !d->mlp_b3)
            return fail(GFMC_EINVAL, "gfmc_problem_create: incomplete MLP parameters");
        if (d->n_sites > 510)
            return fail(GFMC_EINVAL, "gfmc_problem_create: the MLP guide takes at most 510 "
                                     "sites (its layer-1 sums are 32-bit integers at scale "
                                     "2^16 of weights up to 64)");
    }
    if (d->guide != GFMC_GUIDE_NONE && d->guide != GFMC_GUIDE_JASTROW &&
        d->guide != GFMC_GUIDE_RBM && d->guide != GFMC_GUIDE_MLP)
        return fail(GFMC_EINVAL, "gfmc_problem_create: unknown guide kind");
    const int N = d->n_sites, E = d->n_edges, n_off = d->n_hx + d->n_jpm;
    if (d->guide == GFMC_GUIDE_JASTROW &&
        (d->jas_max_deg <= 0 || !d->jas_site_other || !d->jas_site_J || !d->jas_h ||
         (E > 0 && (!d->jas_edge_other || !d->jas_edge_J))))
        return fail(GFMC_EINVAL, "gfmc_problem_create: incomplete Jastrow tables");
    if (d->has_jz && (!d->jz || (!d->dep_off ? false : !d->edge_off)))
        return fail(GFMC_EINVAL, "gfmc_problem_create: J_z tables missing");
    for (int e = 0; e < E; ++e)
        if (d->edge_i[e] < 0 || d->edge_i[e] >= N || d->edge_j[e] < 0 || d->edge_j[e] >= N)
            return fail(GFMC_EINVAL, "gfmc_problem_create: edge index out of range");
    if (d->guide == GFMC_GUIDE_JASTROW) {
        const int D = d->jas_max_deg;
        for (int i = 0; i < N * D; ++i)
            if (d->jas_site_other[i] < 0 || d->jas_site_other[i] >= N)
                return fail(GFMC_EINVAL, "gfmc_problem_create: Jastrow partner out of range");
        for (int i = 0; i < E * 2 * D; ++i)
            if (d->jas_edge_other[i] < 0 || d->jas_edge_other[i] >= N)
                return fail(GFMC_EINVAL, "gfmc_problem_create: Jastrow edge partner out of range");
    }
    if (d->dep_off) {
        if (d->dep_off[0] != 0) return fail(GFMC_EINVAL, "gfmc_problem_create: dep_off[0] != 0");
        for (int s = 0; s < n_off; ++s) {
            if (d->dep_off[s + 1] < d->dep_off[s])
                return fail(GFMC_EINVAL, "gfmc_problem_create: dep_off not monotone");
            for (int t = d->dep_off[s]; t < d->dep_off[s + 1]; ++t)
                if (d->dep_slots[t] < 0 || d->dep_slots[t] >= n_off)
                    return fail(GFMC_EINVAL, "gfmc_problem_create: dependent slot out of range");
            if (d->has_jz) {
                if (d->edge_off[s + 1] < d->edge_off[s])
                    return fail(GFMC_EINVAL, "gfmc_problem_create: edge_off not monotone");
                for (int t = d->edge_off[s]; t < d->edge_off[s + 1]; ++t)
                    if (d->dep_edges[t] < 0 || d->dep_edges[t] >= E)
                        return fail(GFMC_EINVAL, "gfmc_problem_create: dependent edge out of range");
            }
        }
    }

    gfmc_problem *p = new gfmc_problem();
    p->precision = precision;
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            ncu <= 0)
            ncu = 256;
        p->n_cu = ncu;
    }
    DevProblem &P = p->dp;
    memset(&P, 0, sizeof P);
    P.N = N;
    P.n_edges = E;
    P.n_hx = d->n_hx;
    P.n_jpm = d->n_jpm;
    P.n_off = n_off;
    P.has_jz = d->has_jz;
    P.uniform_jz = d->uniform_jz;
    P.has_hz = d->has_hz;
    P.guide = d->guide;
    P.words = (N + 31) / 32;
    P.K = block_k_of(n_off);
    P.Ke = block_k_of(E);
    P.Ks = block_k_of(N);
    P.max_deg = d->guide == GFMC_GUIDE_JASTROW ? d->jas_max_deg : 0;
    P.full = d->dep_off == nullptr;
    P.jz0 = (d->has_jz && E > 0) ? d->jz[0] : 0.0;
    const size_t D = (size_t)P.max_deg;
    HostRecords hr;
    P.fast = build_records(d, hr) ? 1 : 0;
    P.wpb = 1;
    P.xcd = 0;
    P.R = hr.R;
    P.RN = hr.RN;
    P.RE = hr.RE;
    P.flags = d->has_jz && (!d->uniform_jz || !P.fast);
    // the local-field form runs in the FP32 kernels; its zero spin is site N
    // (its kernels take the diagonal from the anti-parallel count: uniform
    // J_z or none, no h_z)
    P.fld = P.fast && hr.fld && precision == GFMC_FP32 && d->guide == GFMC_GUIDE_JASTROW &&
            !d->has_hz && (!d->has_jz || d->uniform_jz);
    P.FD = P.fld ? hr.FD : 0;
    P.zsite = P.fld ? hr.zsite : 0;
    P.n_pad = round_up(N + P.zsite, 16);
    P.e_pad = P.flags ? round_up(E > 0 ? E : 1, 16) : 0;
    P.f_pad = P.fld ? P.n_pad : 0;
    {
        const double jpm0 = E ? d->jpm[0] : 0.0, jpmpm0 = E ? d->jpmpm[0] : 0.0;
        const double hx0 = d->n_hx ? d->hx_half[0] : 0.0;
        p->t64.jpm0 = jpm0;
        p->t64.jpmpm0 = jpmpm0;
        p->t64.hx0 = hx0;
        p->t64.J0 = hr.J0;
        p->t64.h0 = hr.h0;
        p->t32.jpm0 = (float)jpm0;
        p->t32.jpmpm0 = (float)jpmpm0;
        p->t32.hx0 = (float)hx0;
        p->t32.J0 = (float)hr.J0;
        p->t32.h0 = (float)hr.h0;
        p->t32.lJ = (float)(-2.0 * hr.J0 * 1.4426950408889634);
        p->t32.lh = (float)(-2.0 * hr.h0 * 1.4426950408889634);
        p->t64.lJ = -2.0 * hr.J0 * 1.4426950408889634;
        p->t64.lh = -2.0 * hr.h0 * 1.4426950408889634;
    }
    int rc = 0;
#define UP(src, n, dst)                              \
    do {                                             \
        if ((rc = upload(p, src, n, dst)) != 0) {    \
            free_problem(p);                         \
            return rc;                               \
        }                                            \
    } while (0)
#define UP32(src, n, dst)                            \
    do {                                             \
        if ((rc = upload_f32(p, src, n, dst)) != 0) { \
            free_problem(p);                         \
            return rc;                               \
        }                                            \
    } while (0)
    UP(d->edge_i, (size_t)E, &P.edge_i);
    UP(d->edge_j, (size_t)E, &P.edge_j);
    UP(hr.edge_pack.data(), hr.edge_pack.size(), &P.edge_pack);
    if (P.fast) {
        const uint32_t *sr = nullptr, *mr = nullptr;
        UP(hr.slot_rec_lds.data(), hr.slot_rec_lds.size(), &sr);
        UP(hr.mv_rec.data(), hr.mv_rec.size(), &mr);
        P.slot_rec = (const uint4 *)sr;
        P.mv_rec = (const uint4 *)mr;
        P.slot_rec_g[0] = P.slot_rec;
        P.Kg[0] = P.K;
        P.g_ok[0] = 1;
        for (int gi = 1; gi < 3; ++gi) {
            const int G = WAVE >> gi;
            const uint32_t *sg = nullptr;
            UP(hr.slot_rec_g[gi].data(), hr.slot_rec_g[gi].size(), &sg);
            P.slot_rec_g[gi] = (const uint4 *)sg;
            P.Kg[gi] = (n_off + G - 1) / G;
            // every move's records and bond-state neighbours fit the group
            P.g_ok[gi] = hr.R <= G && (!(d->has_jz && d->uniform_jz) || hr.RN <= G);
        }
        UP(hr.mv_nbr.data(), hr.mv_nbr.size(), &P.mv_nbr);
        UP(hr.mv_sites.data(), hr.mv_sites.size(), &P.mv_sites);
        UP(hr.mv_edges.data(), hr.mv_edges.size(), &P.mv_edges);
        if (P.fld) {
            for (int gi = 0; gi < 3; ++gi) {
                const uint32_t *sg = nullptr;
                UP(hr.slot_rf_g[gi].data(), hr.slot_rf_g[gi].size(), &sg);
                P.slot_rf_g[gi] = (const uint2 *)sg;
            }
            for (int gi = 0; gi < 3; ++gi) {
                const uint32_t *mg = nullptr;
                UP(hr.mv_rf_g[gi].data(), hr.mv_rf_g[gi].size(), &mg);
                P.mv_rf_g[gi] = (const uint2 *)mg;
            }
            const uint32_t *fs = nullptr;
            UP(hr.fsite.data(), hr.fsite.size(), &fs);
            P.fsite = (const uint4 *)fs;
            UP(hr.mv_fnb.data(), hr.mv_fnb.size(), &P.mv_fnb);
        }
    }
    if (d->has_jz) UP(d->jz, (size_t)E, &P.jz);
    if (d->has_hz) UP(d->hz, (size_t)N, &P.hz);
    if (d->guide == GFMC_GUIDE_JASTROW) {
        UP(d->jas_site_other, (size_t)N * D, &P.jas_site_other);
        UP(d->jas_edge_other, (size_t)E * 2 * D, &P.jas_edge_other);
    }
    if (d->dep_off) {
        UP(d->dep_off, (size_t)n_off + 1, &P.dep_off);
        UP(d->dep_slots, (size_t)d->dep_off[n_off], &P.dep_slots);
        if (d->has_jz) {
            UP(d->edge_off, (size_t)n_off + 1, &P.edge_off);
            UP(d->dep_edges, (size_t)d->edge_off[n_off], &P.dep_edges);
        }
    }
    // coupling tables in both precisions
    UP(d->jpm, (size_t)E, &p->t64.jpm);
    UP(d->jpmpm, (size_t)E, &p->t64.jpmpm);
    UP32(d->jpm, (size_t)E, &p->t32.jpm);
    UP32(d->jpmpm, (size_t)E, &p->t32.jpmpm);
    p->jpmpm_zero = 1;
    for (int e = 0; e < E; ++e)
        if ((float)d->jpmpm[e] != 0.0f) p->jpmpm_zero = 0;
    if (d->n_hx) {
        UP(d->hx_half, (size_t)N, &p->t64.hx_half);
        UP32(d->hx_half, (size_t)N, &p->t32.hx_half);
    }
    if (d->guide == GFMC_GUIDE_JASTROW) {
        UP(d->jas_h, (size_t)N, &p->t64.jas_h);
        UP32(d->jas_h, (size_t)N, &p->t32.jas_h);
        UP(d->jas_site_J, (size_t)N * D, &p->t64.jas_site_J);
        UP32(d->jas_site_J, (size_t)N * D, &p->t32.jas_site_J);
        UP(d->jas_edge_J, (size_t)E * 2 * D, &p->t64.jas_edge_J);
        UP32(d->jas_edge_J, (size_t)E * 2 * D, &p->t32.jas_edge_J);
    }
    if (d->guide == GFMC_GUIDE_RBM) {
        const int H = d->rbm_hidden, Np = N + (N & 1);
        P.H = H;
        P.w_words = H * Np / 2;
        // W rounded to BF16 (round to nearest even), two per word, rows of Np
        std::vector<uint32_t> wp((size_t)P.w_words, 0u);
        std::vector<float> cs((size_t)N, 0.0f);
        std::vector<double> csd((size_t)N, 0.0);
        for (int h = 0; h < H; ++h)
            for (int k = 0; k < N; ++k) {
                uint32_t bits;
                memcpy(&bits, &d->rbm_W[(size_t)h * N + k], 4);
                const uint32_t b16 = (bits + 0x7fffu + ((bits >> 16) & 1u)) >> 16;
                const size_t idx = (size_t)h * Np + k;
                wp[idx >> 1] |= (idx & 1) ? (b16 << 16) : b16;
                const uint32_t back = b16 << 16;
                float wv;
                memcpy(&wv, &back, 4);
                csd[k] += (double)wv;
            }
        for (int k = 0; k < N; ++k) cs[k] = (float)csd[k];
        const uint32_t *wdev = nullptr;
        UP(wp.data(), wp.size(), &wdev);
        P.rbm_W = wdev;
        UP(d->rbm_a, (size_t)N, &P.rbm_a);
        UP(d->rbm_b, (size_t)H, &P.rbm_b);
        UP(cs.data(), cs.size(), &P.rbm_cs);
        P.rbm_wave_bytes = round_up(P.K * WAVE * 4 + 3 * H * 4 + n_off * 4 + P.n_pad + P.e_pad, 16);
        // the lanes-over-hidden-units kernels: W transposed, [N][Hp] BF16
        P.Hp = round_up(H, WAVE);
        P.wt_bytes = round_up(N * P.Hp * 2, 16);
        P.rbh_wave_bytes = round_up(P.K * WAVE * 4 + (n_off + 4) * 12 + P.n_pad + P.e_pad, 16);
        const int64_t rbh_lds = (int64_t)P.wt_bytes + (int64_t)RBH_WAVES * P.rbh_wave_bytes;
        P.rbh = (P.Hp / WAVE <= RBH_MAX_M && rbh_lds <= 160 * 1024) ? P.Hp / WAVE : 0;
        if (P.rbh) {
            std::vector<uint16_t> wt((size_t)P.wt_bytes / 2, (uint16_t)0);
            for (int h = 0; h < H; ++h)
                for (int k = 0; k < N; ++k) {
                    const size_t idx = (size_t)h * Np + k;
                    wt[(size_t)k * P.Hp + h] = (uint16_t)(wp[idx >> 1] >> ((idx & 1) ? 16 : 0));
                }
            UP(wt.data(), wt.size(), &P.rbm_WT);
            const int zero = 0;
            const int *q = nullptr;
            UP(&zero, (size_t)1, &q);
            P.rbm_queue = const_cast<int *>(q);
            p->rbh_lds_bytes = rbh_lds;
        }
    }
    if (d->guide == GFMC_GUIDE_MLP) {
        const int h1 = d->mlp_h1, h2 = d->mlp_h2;
        P.mlp_h1 = h1;
        P.mlp_h2 = h2;
        std::vector<int32_t> w1i((size_t)N * h1);
        std::vector<uint16_t> w2((size_t)h2 * h1);
        std::vector<double> b1(h1);
        for (int k = 0; k < h1; ++k)
            for (int n = 0; n < N; ++n)
                w1i[(size_t)n * h1 + k] =
                    (int32_t)((double)mlp_l1_value(d->mlp_W1[(size_t)k * N + n]) * 65536.0);
        for (size_t i = 0; i < w2.size(); ++i) w2[i] = mlp_bf16_bits(d->mlp_W2[i]);
        for (int k = 0; k < h1; ++k) b1[k] = (double)mlp_l1_value(d->mlp_b1[k]);
        std::vector<float2> tab(MLP_NT);
        for (int j = 0; j < MLP_NT; ++j) {
            const float t0 = (float)std::tanh((double)j / 256.0);
            const float t1 = (float)std::tanh((double)(j + 1) / 256.0);
            tab[j] = make_float2(t0, t1 - t0);
        }
        UP(w1i.data(), w1i.size(), &P.mlp_W1I);

        UP(w2.data(), w2.size(), &P.mlp_W2);
        UP(b1.data(), b1.size(), &P.mlp_b1);
        UP(d->mlp_b2, (size_t)h2, &P.mlp_b2);
        UP(d->mlp_w3, (size_t)h2, &P.mlp_w3);
        UP(tab.data(), tab.size(), &P.mlp_tab);
        UP(d->mlp_b3, (size_t)1, &P.mlp_b3p);
        P.rbm_wave_bytes = round_up(P.K * WAVE * 4 + P.n_pad + P.e_pad, 16);
        p->mlp = new_mlp_work((n_off + 1 + 15) / 16);
    }
#undef UP
#undef UP32
    const size_t wbytes = (size_t)P.K * WAVE * (precision == GFMC_FP64 ? 8 : 4);
    p->lds_bytes = (int64_t)(wbytes + P.n_pad + P.e_pad + P.f_pad);
    if (d->guide == GFMC_GUIDE_RBM) {
        p->lds_bytes = (int64_t)P.w_words * 4 + (int64_t)RBM_WAVES * P.rbm_wave_bytes;
        if (p->lds_bytes > 160 * 1024) {
            free_problem(p);
            return fail(GFMC_EUNSUP, "gfmc_problem_create: RBM weights plus walker state "
                                     "exceed the 160 KiB of LDS per workgroup");
        }
        HIP_TRY(hipFuncSetAttribute((const void *)k_propagate_rbm,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)p->lds_bytes));
        HIP_TRY(hipFuncSetAttribute((const void *)k_rates_rbm,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)p->lds_bytes));
        if (P.rbh) {
            int rc2 = rbh_set_lds(P.rbh, (int)p->rbh_lds_bytes);
            if (rc2) {
                free_problem(p);
                return rc2;
            }
        }
    } else if (d->guide == GFMC_GUIDE_MLP) {
        p->lds_bytes = (int64_t)MLP_SW * P.rbm_wave_bytes;
        if (p->lds_bytes > 64 * 1024) {
            free_problem(p);
            return fail(GFMC_EUNSUP, "gfmc_problem_create: MLP walker state exceeds 64 KiB of "
                                     "LDS per workgroup");
        }
    } else if (p->lds_bytes > 64 * 1024) {
        free_problem(p);
        return fail(GFMC_EUNSUP,
                    "gfmc_problem_create: walker state exceeds 64 KiB of LDS per wavefront");
    }
    *out = p;
    return 0;
}

int gfmc_problem_destroy(gfmc_problem *p) {
    if (p) free_problem(p);
    return 0;
}

int gfmc_rbm_set_tied(gfmc_problem *p, int32_t n_filters, int32_t n_T, int32_t Q,
                      int32_t n_cls, int32_t pair_code, int32_t n_pair_chunks,
                      int32_t n_single_chunks, const int32_t *lane_off,
                      const int32_t *lane_unit, const int32_t *slot_code,
                      const int32_t *site_code, const int32_t *q_site, const int32_t *q_pair,
                      void *stream) {
    const char *fn = "gfmc_rbm_set_tied";
    if (!p) return fail(GFMC_EINVAL, std::string(fn) + ": NULL problem");
    DevProblem &P = p->dp;
    if (P.guide != GFMC_GUIDE_RBM)
        return fail(GFMC_EINVAL, std::string(fn) + ": the problem's guide is not an RBM");
    if (P.rbt) return fail(GFMC_EINVAL, std::string(fn) + ": the tables are already set");
    if (!P.rbh)
        return fail(GFMC_EUNSUP, std::string(fn) + ": hidden layer wider than the "
                                                   "lanes-over-units kernels take");
    if (P.n_hx != 0 || !p->jpmpm_zero || P.n_off != P.n_edges)
        return fail(GFMC_EUNSUP, std::string(fn) + ": the tied kernel takes pair moves of "
                                                   "anti-parallel spins only (no h_x, J_pmpm = 0)");
    if (n_filters <= 0 || n_T <= 0 || (int64_t)n_filters * n_T != P.H || Q <= 0 ||
        Q >= (1 << 20) || n_cls <= 0 || n_cls >= (1 << 10))
        return fail(GFMC_EINVAL, std::string(fn) + ": bad sizes");
    if (!lane_off || !lane_unit || !slot_code || !site_code || !q_site || !q_pair)
        return fail(GFMC_EINVAL, std::string(fn) + ": NULL table");
    const int cp = n_pair_chunks, cs = n_single_chunks;
    if (cp < 0 || cs < 0 || (int64_t)(2 * cp + cs) * WAVE < P.H)
        return fail(GFMC_EINVAL, std::string(fn) + ": the lane chunks hold fewer units than H");
    if (!rbt_shape_ok(cp, cs))
        return fail(GFMC_EUNSUP, std::string(fn) + ": no tied kernel for this chunk shape");
    const int N = P.N, n_off = P.n_off;
    const int64_t FQ = (int64_t)n_filters * Q;
    const int64_t raw = (2 * (int64_t)n_cls + 1) * FQ + Q;
    const int64_t words = (2 * raw + 3) / 4 * 4;
    if (pair_code < 0 || pair_code >= raw || 8 * raw >= INT32_MAX)
        return fail(GFMC_EINVAL, std::string(fn) + ": bad pair code");
    // every hidden unit sits in exactly one lane register
    {
        std::vector<int> seen(P.H, 0);
        for (int64_t t = 0; t < (int64_t)(2 * cp + cs) * WAVE; ++t) {
            const int h = lane_unit[t];
            if (h < -1 || h >= P.H || (h >= 0 && seen[h]++))
                return fail(GFMC_EINVAL, std::string(fn) + ": lane units are not a partition "
                                                           "of the hidden units");
        }
        for (int h = 0; h < P.H; ++h)
            if (!seen[h])
                return fail(GFMC_EINVAL, std::string(fn) + ": a hidden unit has no lane");
    }
    // every table read of the kernel stays inside the table: theta reads
    // 2 n_cls FQ + code(k) + off, factor reads (2 cls + sg) FQ + code(i) + off
    int64_t kmin = INT64_MAX, kmax = INT64_MIN, smin = INT64_MAX, smax = INT64_MIN;
    for (int k = 0; k < N; ++k) {
        if (site_code[k] < 0 || site_code[k] >= Q)
            return fail(GFMC_EINVAL, std::string(fn) + ": site code out of range");
        kmin = std::min<int64_t>(kmin, site_code[k]);
        kmax = std::max<int64_t>(kmax, site_code[k]);
    }
    for (int s = 0; s < n_off; ++s) {
        const int code = slot_code[s] & 0xfffff, cls = (slot_code[s] >> 20) & 0x3ff;
        if (slot_code[s] < 0 || (slot_code[s] >> 31) || cls >= n_cls)
            return fail(GFMC_EINVAL, std::string(fn) + ": slot class out of range");
        smin = std::min<int64_t>(smin, 2 * (int64_t)cls * FQ + code);
        smax = std::max<int64_t>(smax, (2 * (int64_t)cls + 1) * FQ + code);
    }
    for (int64_t t = 0; t < (int64_t)(cp + cs) * WAVE; ++t) {
        const int64_t v = lane_off[t];
        if (2 * n_cls * FQ + kmin + v < 0 || 2 * n_cls * FQ + kmax + v >= raw || smin + v < 0 ||
            smax + v >= raw)
            return fail(GFMC_EINVAL, std::string(fn) + ": unit offset reads outside the tables");
    }
    for (int q = 0; q < Q; ++q)
        if (q_site[q] < 0 || q_site[q] >= N)
            return fail(GFMC_EINVAL, std::string(fn) + ": entry site out of range");
    for (int64_t t = 0; t < (int64_t)n_cls * Q; ++t)
        if (q_pair[t] < -1 || q_pair[t] >= N)
            return fail(GFMC_EINVAL, std::string(fn) + ": entry partner out of range");
    int waves = 16;
    while (waves >= 4 && words * 4 + (int64_t)waves * P.rbh_wave_bytes > 160 * 1024) waves /= 2;
    if (waves < 4)
        return fail(GFMC_EUNSUP, std::string(fn) + ": tables plus walker state exceed the LDS");
    const int64_t lds = words * 4 + (int64_t)waves * P.rbh_wave_bytes;
    int rc = 0;
    const int *d_unit = nullptr, *d_hid = nullptr, *d_scode = nullptr, *d_kcode = nullptr,
              *d_qsite = nullptr, *d_qpair = nullptr;
    if ((rc = upload(p, lane_off, (size_t)(cp + cs) * WAVE, &d_unit)) ||
        (rc = upload(p, lane_unit, (size_t)(2 * cp + cs) * WAVE, &d_hid)) ||
        (rc = upload(p, slot_code, (size_t)n_off, &d_scode)) ||
        (rc = upload(p, site_code, (size_t)N, &d_kcode)) ||
        (rc = upload(p, q_site, (size_t)Q, &d_qsite)) ||
        (rc = upload(p, q_pair, (size_t)n_cls * Q, &d_qpair)))
        return rc;
    void *tab = nullptr, *rawp = nullptr;
    HIP_TRY(hipMalloc(&tab, (size_t)words * 4));
    p->allocs.push_back(tab);
    HIP_TRY(hipMemset(tab, 0, (size_t)words * 4));
    HIP_TRY(hipMalloc(&rawp, (size_t)raw * 4));
    p->allocs.push_back(rawp);
    // a single chunk of at most 16 units is packed four slots to a chunk
    bool pk = cs == 1;
    for (int l = 16; pk && l < WAVE; ++l) pk = lane_unit[(int64_t)2 * cp * WAVE + l] < 0;
    if ((rc = RBT_SWITCH(cp, cs, pk, rbt_set_lds_m, (int)lds))) return rc;
    P.rbt_pk = pk ? 1 : 0;
    P.rbt_cp = cp;
    P.rbt_cs = cs;
    P.rbt_D = pair_code;
    P.rbt_F = n_filters;
    P.rbt_nT = n_T;
    P.rbt_Q = Q;
    P.rbt_ncls = n_cls;
    P.rbt_words = (int)words;
    P.rbt_tab = (const float *)tab;
    P.rbt_raw = (float *)rawp;
    P.rbt_unit = d_unit;
    P.rbt_hid = d_hid;
    P.rbt_scode = d_scode;
    P.rbt_kcode = d_kcode;
    P.rbt_qsite = d_qsite;
    P.rbt_qpair = d_qpair;
    p->rbt_lds_bytes = lds;
    p->rbt_waves = waves;
    rbt_build_tables(p, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    P.rbt = 1;
    return 0;
}

int gfmc_rbm_untie(gfmc_problem *p) {
    if (!p) return fail(GFMC_EINVAL, "gfmc_rbm_untie: NULL problem");
    p->dp.rbt = 0;
    return 0;
}

int gfmc_problem_set_launch(gfmc_problem *p, int32_t waves_per_block, int32_t xcd_remap) {
    if (!p) return fail(GFMC_EINVAL, "gfmc_problem_set_launch: NULL problem");
    if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4)
        return fail(GFMC_EINVAL, "gfmc_problem_set_launch: waves_per_block must be 1, 2 or 4");
    if (xcd_remap != 0 && xcd_remap != 1)
        return fail(GFMC_EINVAL, "gfmc_problem_set_launch: xcd_remap must be 0 or 1");
    if (waves_per_block > 1 && !p->dp.fast)
        return fail(GFMC_EUNSUP, "gfmc_problem_set_launch: this problem's propagation runs one "
                                 "wavefront per workgroup (generic path)");
    p->dp.wpb = waves_per_block;
    p->dp.xcd = xcd_remap;
    return 0;
}

int gfmc_problem_set_lanes(gfmc_problem *p, int32_t lanes) {
    if (!p) return fail(GFMC_EINVAL, "gfmc_problem_set_lanes: NULL problem");
    if (lanes != 64 && lanes != 32 && lanes != 16)
        return fail(GFMC_EINVAL, "gfmc_problem_set_lanes: lanes must be 64, 32 or 16");
    if (lanes != 64 && (p->dp.guide == GFMC_GUIDE_RBM || p->dp.guide == GFMC_GUIDE_MLP ||
                        !p->dp.fast ||
                        !p->dp.g_ok[gidx_of(lanes)] || !lanes_lds_fits(p, lanes)))
        return fail(GFMC_EUNSUP, "gfmc_problem_set_lanes: this problem runs one walker per "
                                 "wavefront only (record path, moves within the group, LDS)");
    p->lanes = lanes;
    return 0;
}

int gfmc_problem_set_method(gfmc_problem *p, int32_t method, double lam) {
    if (!p) return fail(GFMC_EINVAL, "gfmc_problem_set_method: NULL problem");
    if (method != GFMC_METHOD_CONT && method != GFMC_METHOD_SINGLE &&
        method != GFMC_METHOD_MULTI)
        return fail(GFMC_EINVAL, "gfmc_problem_set_method: unknown method");
    if (method != GFMC_METHOD_CONT) {
        if (!std::isfinite(lam))
            return fail(GFMC_EINVAL, "gfmc_problem_set_method: lam must be finite");
        if (p->dp.guide == GFMC_GUIDE_RBM || p->dp.guide == GFMC_GUIDE_MLP)
            return fail(GFMC_EUNSUP, "gfmc_problem_set_method: the network guides propagate "
                                     "in continuous time only");
    }
    p->dp.method = method;
    p->dp.lam = method == GFMC_METHOD_CONT ? 0.0 : lam;
    return 0;
}

int gfmc_problem_get_info(const gfmc_problem *p, gfmc_problem_info *out) {
    if (!p || !out) return fail(GFMC_EINVAL, "gfmc_problem_get_info: NULL argument");
    out->n_sites = p->dp.N;
    out->n_words = p->dp.words;
    out->n_off = p->dp.n_off;
    out->m_total = p->dp.n_off + 1;
    out->precision = p->precision;
    out->guide = p->dp.guide;
    out->block_k = p->dp.K;
    out->full_recompute = p->dp.full;
    out->lds_bytes = p->lds_bytes;
    out->lanes = p->lanes;
    out->lanes_ok = 0;
    for (int gi = 0; gi < 3; ++gi)
        if (gi == 0 || (p->dp.guide != GFMC_GUIDE_MLP && p->dp.fast && p->dp.g_ok[gi] &&
                        lanes_lds_fits(p, WAVE >> gi)))
            out->lanes_ok |= WAVE >> gi;
    return 0;
}

static unsigned grid1d(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

int gfmc_init_walkers_sz(int32_t n_sites, int64_t nw, int32_t n_up, uint64_t seed,
                         uint32_t *spins, void *stream) {
    if (n_sites <= 0 || nw < 0 || n_up < 0 || n_up > n_sites || (nw > 0 && !spins))
        return fail(GFMC_EINVAL, "gfmc_init_walkers_sz: bad argument");
    if (nw == 0) return 0;
    if (nw > 0x7fffffffLL) return fail(GFMC_EINVAL, "gfmc_init_walkers_sz: too many walkers");
    const int words = (n_sites + 31) / 32;
    size_t lds = (size_t)round_up(n_sites, 16);
    hipLaunchKernelGGL(k_init_sz, dim3((unsigned)nw), dim3(WAVE), lds, (hipStream_t)stream,
                       (int)n_sites, words, nw, (int)n_up, (uint32_t)seed,
                       (uint32_t)(seed >> 32), spins);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_init_walkers_random(int32_t n_sites, int64_t nw, uint64_t seed, uint32_t *spins,
                             void *stream) {
    if (n_sites <= 0 || nw < 0 || (nw > 0 && !spins))
        return fail(GFMC_EINVAL, "gfmc_init_walkers_random: bad argument");
    if (nw == 0) return 0;
    const int words = (n_sites + 31) / 32;
    hipLaunchKernelGGL(k_init_random, dim3(grid1d(nw * words, 256)), dim3(256), 0,
                       (hipStream_t)stream, (int)n_sites, words, nw, (uint32_t)seed,
                       (uint32_t)(seed >> 32), spins);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_pack(int32_t n_sites, int64_t nw, const double *config, uint32_t *spins, void *stream) {
    if (n_sites <= 0 || nw < 0 || (nw > 0 && (!config || !spins)))
        return fail(GFMC_EINVAL, "gfmc_pack: bad argument");
    if (nw == 0) return 0;
    const int words = (n_sites + 31) / 32;
    hipLaunchKernelGGL(k_pack, dim3(grid1d(nw * words, 256)), dim3(256), 0, (hipStream_t)stream,
                       (int)n_sites, words, nw, config, spins);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_unpack(int32_t n_sites, int64_t nw, const uint32_t *spins, double *config,
                void *stream) {
    if (n_sites <= 0 || nw < 0 || (nw > 0 && (!config || !spins)))
        return fail(GFMC_EINVAL, "gfmc_unpack: bad argument");
    if (nw == 0) return 0;
    const int words = (n_sites + 31) / 32;
    hipLaunchKernelGGL(k_unpack, dim3(grid1d(nw * n_sites, 256)), dim3(256), 0,
                       (hipStream_t)stream, (int)n_sites, words, nw, spins, config);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int check_walkers(const char *fn, const gfmc_problem *p, int64_t nw) {
    if (!p) return fail(GFMC_EINVAL, std::string(fn) + ": NULL problem");
    if (nw < 0 || nw > 0x7fffffffLL)
        return fail(GFMC_EINVAL, std::string(fn) + ": n_walkers out of range");
    return 0;
}

int gfmc_rates(const gfmc_problem *p, int64_t nw, const uint32_t *spins, double *rates,
               double *e_loc, double *log_diff, void *stream) {
    int rc = check_walkers("gfmc_rates", p, nw);
    if (rc) return rc;
    if (nw == 0) return 0;
    if (!spins) return fail(GFMC_EINVAL, "gfmc_rates: NULL spins");
    if (p->dp.guide == GFMC_GUIDE_MLP)
        return mlp_rates(p, nw, spins, rates, e_loc, log_diff, nullptr, (hipStream_t)stream);
    if (p->dp.guide == GFMC_GUIDE_RBM && p->dp.rbh)
        RBH_SWITCH(p->dp.rbh, rbh_rates_m, p, nw, spins, rates, e_loc, log_diff,
                   (double *)nullptr, (hipStream_t)stream);
    else if (p->dp.guide == GFMC_GUIDE_RBM)
        hipLaunchKernelGGL(k_rates_rbm, dim3(rbm_grid(p, nw)), dim3(RBM_WAVES * WAVE),
                           (size_t)p->lds_bytes, (hipStream_t)stream, p->dp, p->t32, nw, spins,
                           rates, e_loc, log_diff, (double *)nullptr);
    else if (p->precision == GFMC_FP64)
        launch_rates(p, p->t64, nw, spins, rates, e_loc, log_diff, (hipStream_t)stream);
    else
        launch_rates(p, p->t32, nw, spins, rates, e_loc, log_diff, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_log_psi(const gfmc_problem *p, int64_t nw, const uint32_t *spins, double *out,
                 void *stream) {
    int rc = check_walkers("gfmc_log_psi", p, nw);
    if (rc) return rc;
    if (nw == 0) return 0;
    if (!spins || !out) return fail(GFMC_EINVAL, "gfmc_log_psi: NULL buffer");
    if (p->dp.guide == GFMC_GUIDE_MLP)
        return mlp_rates(p, nw, spins, nullptr, nullptr, nullptr, out, (hipStream_t)stream);
    if (p->dp.guide == GFMC_GUIDE_RBM && p->dp.rbh)
        RBH_SWITCH(p->dp.rbh, rbh_rates_m, p, nw, spins, (double *)nullptr, (double *)nullptr,
                   (double *)nullptr, out, (hipStream_t)stream);
    else if (p->dp.guide == GFMC_GUIDE_RBM)
        hipLaunchKernelGGL(k_rates_rbm, dim3(rbm_grid(p, nw)), dim3(RBM_WAVES * WAVE),
                           (size_t)p->lds_bytes, (hipStream_t)stream, p->dp, p->t32, nw, spins,
                           (double *)nullptr, (double *)nullptr, (double *)nullptr, out);
    else
    hipLaunchKernelGGL(k_log_psi, dim3(grid1d(nw, 64)), dim3(64), 0, (hipStream_t)stream, p->dp,
                       p->t64, nw, spins, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_propagate(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                   double *e_loc, double beta, double alw, uint64_t seed, uint32_t step,
                   uint32_t tag, int32_t *evals, void *stream) {
    int rc = check_walkers("gfmc_propagate", p, nw);
    if (rc) return rc;
    if (nw == 0) return 0;
    if (!spins || !log_w || !e_loc) return fail(GFMC_EINVAL, "gfmc_propagate: NULL buffer");
    if (!(beta > 0) || !std::isfinite(beta))
        return fail(GFMC_EINVAL, "gfmc_propagate: beta must be finite and > 0");
    if (p->dp.method != GFMC_METHOD_CONT && (beta != std::floor(beta) || beta > 1e9))
        return fail(GFMC_EINVAL, "gfmc_propagate: the discrete methods take an integer "
                                 "number of sub-steps");
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (p->dp.guide == GFMC_GUIDE_MLP)
        return mlp_propagate(p, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag, evals,
                             (hipStream_t)stream);
    if (p->dp.guide == GFMC_GUIDE_RBM && p->dp.rbt)
        RBT_SWITCH(p->dp.rbt_cp, p->dp.rbt_cs, p->dp.rbt_pk, rbt_launch_m, p, nw, spins, log_w,
                   e_loc, beta, alw, k0, k1, step, tag, evals, (hipStream_t)stream);
    else if (p->dp.guide == GFMC_GUIDE_RBM && p->dp.rbh)
        RBH_SWITCH(p->dp.rbh, rbh_launch_m, p, nw, spins, log_w, e_loc, beta, alw, k0, k1, step,
                   tag, evals, (hipStream_t)stream);
    else if (p->dp.guide == GFMC_GUIDE_RBM)
        hipLaunchKernelGGL(k_propagate_rbm, dim3(rbm_grid(p, nw)), dim3(RBM_WAVES * WAVE),
                           (size_t)p->lds_bytes, (hipStream_t)stream, p->dp, p->t32, nw, spins,
                           log_w, e_loc, beta, alw, k0, k1, step, tag, evals, g_step_base);
    else if (p->precision == GFMC_FP64)
        launch_propagate(p, p->t64, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag,
                         evals, (hipStream_t)stream);
    else
        launch_propagate(p, p->t32, nw, spins, log_w, e_loc, beta, alw, k0, k1, step, tag,
                         evals, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_reconfigure(int64_t nw, const double *log_w, const double *e_loc, uint64_t seed,
                     uint32_t step, uint32_t tag, const double *psi_opt, int32_t *parents,
                     double *out2, double *work, void *stream) {
    if (nw <= 0 || nw > 0x7fffffffLL)
        return fail(GFMC_EINVAL, "gfmc_reconfigure: n_walkers must be in [1, 2^31)");
    if (!log_w || !e_loc || !parents || !out2 || !work)
        return fail(GFMC_EINVAL, "gfmc_reconfigure: NULL buffer");
    double psi = psi_opt ? *psi_opt : 0.0;
    if (psi_opt && !(psi >= 0.0 && psi < 1.0))
        return fail(GFMC_EINVAL, "gfmc_reconfigure: psi must lie in [0, 1)");
    RcSegs S;
    S.nseg = 0;
    S.n = nw;
    S.w = nullptr;
    S.b = nullptr;
    launch_reconfigure(S, nw, (unsigned)((nw + RB - 1) / RB), log_w, e_loc, seed, step, tag,
                       psi_opt ? 1 : 0, psi, parents, out2, work, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

int64_t gfmc_reconfigure_work_doubles(int64_t n_walkers) {
    if (n_walkers <= 0) return 0;
    return n_walkers + 7 * ((n_walkers + RB - 1) / RB) + 1;
}

// ---------------------------------------------------------------------------
// reconfiguration that allocates nothing: its scratch (the prefix sum and the
// per-block partials of gfmc_reconfigure's workspace) is a file-scope device
// array, one copy per GPU, resident until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in
// symbol): the device address comes from hipGetSymbolAddress, once per device
// with that device current, and is kept per device.
// ---------------------------------------------------------------------------
#define RC_STATIC_DOUBLES                                                      \
    (GFMC_RECONFIGURE_STATIC_MAX_WALKERS +                                     \
     7 * ((GFMC_RECONFIGURE_STATIC_MAX_WALKERS + RB - 1) / RB) + 1)
static __device__ __align__(256) double scratch[RC_STATIC_DOUBLES];

#define RC_STATIC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double *g_scratch_dev[RC_STATIC_MAX_DEVICES];

static int rc_static_scratch(double **out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= RC_STATIC_MAX_DEVICES)
        return fail(GFMC_EUNSUP, "gfmc_reconfigure_static: device ordinal beyond " +
                                     std::to_string(RC_STATIC_MAX_DEVICES - 1));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(GFMC_EHIP, std::string("gfmc_reconfigure_static: hipGetSymbolAddress: ") +
                                       hipGetErrorString(e));
        g_scratch_dev[dev] = (double *)ptr;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

int gfmc_reconfigure_static(int64_t nw, const double *log_w, const double *e_loc,
                            uint64_t seed, uint32_t step, uint32_t tag, const double *psi_opt,
                            int32_t *parents, double *out2, void *stream) {
    if (nw <= 0 || nw > GFMC_RECONFIGURE_STATIC_MAX_WALKERS)
        return fail(GFMC_EINVAL, "gfmc_reconfigure_static: n_walkers must be in [1, " +
                                     std::to_string(GFMC_RECONFIGURE_STATIC_MAX_WALKERS) +
                                     "] (the size of its static scratch); use "
                                     "gfmc_reconfigure with a workspace for more");
    if (!log_w || !e_loc || !parents || !out2)
        return fail(GFMC_EINVAL, "gfmc_reconfigure_static: NULL buffer");
    double psi = psi_opt ? *psi_opt : 0.0;
    if (psi_opt && !(psi >= 0.0 && psi < 1.0))
        return fail(GFMC_EINVAL, "gfmc_reconfigure_static: psi must lie in [0, 1)");
    double *work = nullptr;
    int rc = rc_static_scratch(&work);
    if (rc) return rc;
    // every scratch entry the kernels read is written earlier in the same
    // call (block partials, then the prefix sum), so nothing needs clearing
    RcSegs S;
    S.nseg = 0;
    S.n = nw;
    S.w = nullptr;
    S.b = nullptr;
    launch_reconfigure(S, nw, (unsigned)((nw + RB - 1) / RB), log_w, e_loc, seed, step, tag,
                       psi_opt ? 1 : 0, psi, parents, out2, work, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_gather(int32_t n_sites, int64_t nw, const uint32_t *spins_in, const double *e_loc_in,
                const int32_t *parents, uint32_t *spins_out, double *e_loc_out,
                double *log_w_out, void *stream) {
    if (n_sites <= 0 || nw < 0) return fail(GFMC_EINVAL, "gfmc_gather: bad size");
    if (nw == 0) return 0;
    if (!spins_in || !parents || !spins_out || (e_loc_out && !e_loc_in))
        return fail(GFMC_EINVAL, "gfmc_gather: NULL buffer");
    const int words = (n_sites + 31) / 32;
    hipLaunchKernelGGL(k_gather, dim3(grid1d(nw * words, 256)), dim3(256), 0,
                       (hipStream_t)stream, words, nw, spins_in, e_loc_in, parents, spins_out,
                       e_loc_out, log_w_out);
    HIP_TRY(hipGetLastError());
    return 0;
}

// the tied-RBM kernel's rows (in 4-byte units): slot weights, theta, spins
// (the diagonal then comes from the anti-parallel count alone: uniform J_z)
static bool rbt_carries(const gfmc_problem *p) {
    return p->dp.guide == GFMC_GUIDE_RBM && p->dp.rbt && (!p->dp.has_jz || p->dp.uniform_jz);
}
static int64_t rbt_theta_off(const gfmc_problem *p) { return (int64_t)p->dp.K * WAVE; }
static int64_t rbt_spins_off(const gfmc_problem *p) {
    return rbt_theta_off(p) + 2 * (int64_t)(2 * p->dp.rbt_cp + p->dp.rbt_cs) * WAVE;
}

// elements of one carried row: the largest weight array over the usable
// lane counts (0 when the problem's path carries nothing)
static int64_t carry_weights(const gfmc_problem *p) {
    if (!p->dp.fast || p->dp.flags) return 0;
    int64_t row = 0;
    for (int gi = 0; gi < 3; ++gi) row = std::max(row, (int64_t)p->dp.Kg[gi] * (WAVE >> gi));
    return row;
}

// ... followed, for the local-field form, by the walker's fields (f_pad
// bytes, a multiple of 16), then by the spins the walker ended its step with
// (the words rounded up to 16 bytes), against which the next step checks the
// spins it loads
static int64_t carry_spins_off(const gfmc_problem *p) {
    const int64_t el = p->precision == GFMC_FP64 ? 8 : 4;
    return carry_weights(p) + (p->dp.fld ? p->dp.f_pad / el : 0);
}

static int64_t carry_row(const gfmc_problem *p) {
    if (rbt_carries(p)) return rbt_spins_off(p) + ((int64_t)p->dp.words * 4 + 15) / 16 * 4;
    const int64_t w = carry_weights(p);
    if (w == 0) return 0;
    const int64_t el = p->precision == GFMC_FP64 ? 8 : 4;
    return carry_spins_off(p) + ((int64_t)p->dp.words * 4 + 15) / 16 * 16 / el;
}

static size_t carry_align(size_t b) { return (b + 255) / 256 * 256; }

static size_t carry_el(const gfmc_problem *p) {
    return rbt_carries(p) ? 4 : (p->precision == GFMC_FP64 ? 8 : 4);
}

static int64_t carry_bytes_of(const gfmc_problem *p, int64_t nw) {
    const int64_t row = carry_row(p);
    if (row == 0 || nw <= 0) return 0;
    const size_t el = carry_el(p);
    return (int64_t)(256 + 2 * carry_align((size_t)nw * row * el) +
                     2 * carry_align((size_t)nw * 4));
}

static CarryDev carry_dev(const gfmc_problem *p, int64_t nw, void *carry,
                          const int32_t *parents) {
    CarryDev c{};
    const int64_t row = carry_row(p);
    if (row == 0) return c;
    const size_t el = carry_el(p);
    unsigned char *b = (unsigned char *)carry;
    c.state = (int32_t *)b;
    size_t o = 256;
    const size_t wb = carry_align((size_t)nw * row * el), ab = carry_align((size_t)nw * 4);
    c.w[0] = b + o;
    c.w[1] = b + o + wb;
    c.anti[0] = (int32_t *)(b + o + 2 * wb);
    c.anti[1] = (int32_t *)(b + o + 2 * wb + ab);
    c.parents = parents;
    c.f_off = rbt_carries(p) ? rbt_theta_off(p) : (p->dp.fld ? carry_weights(p) : 0);
    c.s_off = rbt_carries(p) ? rbt_spins_off(p) : carry_spins_off(p);
    c.row = row;
    c.on = 1;
    return c;
}

// the step loop, with the diagonal forward-walking record after each
// resampling when dg is non-null (gfmc_diag.inl)
static int run_steps_impl(const char *fn, const gfmc_problem *p, int64_t nw, uint32_t *spins,
                          double *log_w, double *e_loc, uint32_t *spins_b, double *e_loc_b,
                          int32_t *parents, double *work, int32_t n_steps, double beta,
                          double alw, uint64_t seed, uint32_t step0, uint32_t tag_prop,
                          uint32_t tag_reconf, double *series, int32_t *evals,
                          const gfmc_diag *dg, void *ring, double *dseries, double *dwork,
                          void *stream, void *carry = nullptr) {
    int rc = check_walkers(fn, p, nw);
    if (rc) return rc;
    if (nw == 0 || n_steps < 0) return fail(GFMC_EINVAL, std::string(fn) + ": bad size");
    if (!spins || !log_w || !e_loc || !spins_b || !e_loc_b || !parents || !work ||
        (n_steps > 0 && !series))
        return fail(GFMC_EINVAL, std::string(fn) + ": NULL buffer");
    if (dg) {
        rc = diag_check(fn, dg, p, ring, dwork);
        if (rc) return rc;
        if (n_steps > 0 && !dseries)
            return fail(GFMC_EINVAL, std::string(fn) + ": NULL diagonal series");
    }
    const int words = p->dp.words;
    uint32_t *sa = spins, *sb = spins_b;
    double *ea = e_loc, *eb = e_loc_b;
    CarryDev cy{};
    if (carry) cy = carry_dev(p, nw, carry, parents);
    for (int32_t t = 0; t < n_steps; ++t) {
        g_carry = cy.on ? &cy : nullptr;
        rc = gfmc_propagate(p, nw, sa, log_w, ea, beta, alw, seed, step0 + (uint32_t)t,
                            tag_prop, evals, stream);
        g_carry = nullptr;
        if (rc) return rc;
        rc = gfmc_reconfigure(nw, log_w, ea, seed, step0 + (uint32_t)t, tag_reconf, nullptr,
                              parents, series + 2 * (int64_t)t, work, stream);
        if (rc) return rc;
        rc = gfmc_gather(p->dp.N, nw, sa, ea, parents, sb, eb, log_w, stream);
        if (rc) return rc;
        if (dg) {
            rc = diag_record_impl(dg, nw, sb, parents, ring,
                                  dseries + (int64_t)t * diag_row_doubles(dg), dwork,
                                  (hipStream_t)stream);
            if (rc) return rc;
        }
        uint32_t *ts = sa;
        sa = sb;
        sb = ts;
        double *te = ea;
        ea = eb;
        eb = te;
    }
    if (sa != spins) {
        HIP_TRY(hipMemcpyAsync(spins, sa, (size_t)nw * words * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, (hipStream_t)stream));
        HIP_TRY(hipMemcpyAsync(e_loc, ea, (size_t)nw * sizeof(double), hipMemcpyDeviceToDevice,
                               (hipStream_t)stream));
    }
    return 0;
}

int gfmc_run_steps(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                   double *e_loc, uint32_t *spins_b, double *e_loc_b, int32_t *parents,
                   double *work, int32_t n_steps, double beta, double alw, uint64_t seed,
                   uint32_t step0, uint32_t tag_prop, uint32_t tag_reconf, double *series,
                   int32_t *evals, void *stream) {
    return run_steps_impl("gfmc_run_steps", p, nw, spins, log_w, e_loc, spins_b, e_loc_b,
                          parents, work, n_steps, beta, alw, seed, step0, tag_prop, tag_reconf,
                          series, evals, nullptr, nullptr, nullptr, nullptr, stream);
}

int64_t gfmc_carry_bytes(const gfmc_problem *p, int64_t n_walkers) {
    if (!p) return 0;
    return carry_bytes_of(p, n_walkers);
}

int gfmc_propagate_carry(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                         double *e_loc, double beta, double alw, uint64_t seed, uint32_t step,
                         uint32_t tag, int32_t *evals, const int32_t *parents, void *carry,
                         void *stream) {
    if (carry && ((uintptr_t)carry & 255))
        return fail(GFMC_EINVAL, "gfmc_propagate_carry: carry must be 256-byte aligned");
    if (carry && !parents) return fail(GFMC_EINVAL, "gfmc_propagate_carry: NULL parents");
    int rc = check_walkers("gfmc_propagate_carry", p, nw);
    if (rc) return rc;
    CarryDev cy{};
    if (carry && nw > 0) cy = carry_dev(p, nw, carry, parents);
    g_carry = cy.on ? &cy : nullptr;
    rc = gfmc_propagate(p, nw, spins, log_w, e_loc, beta, alw, seed, step, tag, evals, stream);
    g_carry = nullptr;
    return rc;
}

int gfmc_run_steps_carry(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                         double *e_loc, uint32_t *spins_b, double *e_loc_b, int32_t *parents,
                         double *work, int32_t n_steps, double beta, double alw, uint64_t seed,
                         uint32_t step0, uint32_t tag_prop, uint32_t tag_reconf, double *series,
                         int32_t *evals, const gfmc_diag *d, void *ring, double *diag_series,
                         double *diag_work, void *carry, void *stream) {
    if (carry && ((uintptr_t)carry & 255))
        return fail(GFMC_EINVAL, "gfmc_run_steps_carry: carry must be 256-byte aligned");
    return run_steps_impl("gfmc_run_steps_carry", p, nw, spins, log_w, e_loc, spins_b, e_loc_b,
                          parents, work, n_steps, beta, alw, seed, step0, tag_prop, tag_reconf,
                          series, evals, d, ring, diag_series, diag_work, stream, carry);
}

int gfmc_sample_index(int32_t precision, int32_t n, const double *weights, int32_t n_draws,
                      const double *uniforms, int32_t *out, void *stream) {
    if (n <= 0 || n_draws < 0 || !weights || (n_draws > 0 && (!uniforms || !out)))
        return fail(GFMC_EINVAL, "gfmc_sample_index: bad argument");
    if (n_draws == 0) return 0;
    const int K = (n + WAVE - 1) / WAVE;
    if (precision == GFMC_FP64)
        hipLaunchKernelGGL(k_sample<double>, dim3(1), dim3(WAVE), (size_t)K * WAVE * 8,
                           (hipStream_t)stream, n, weights, n_draws, uniforms, out);
    else
        hipLaunchKernelGGL(k_sample<float>, dim3(1), dim3(WAVE), (size_t)K * WAVE * 4,
                           (hipStream_t)stream, n, weights, n_draws, uniforms, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_wave_scan(int32_t precision, const double *in, double *out, void *stream) {
    if (!in || !out) return fail(GFMC_EINVAL, "gfmc_wave_scan: NULL buffer");
    if (precision == GFMC_FP64)
        hipLaunchKernelGGL(k_wave_scan<double>, dim3(1), dim3(WAVE), 0, (hipStream_t)stream, in,
                           out);
    else
        hipLaunchKernelGGL(k_wave_scan<float>, dim3(1), dim3(WAVE), 0, (hipStream_t)stream, in,
                           out);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gfmc_philox_host(int32_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    if (n < 0 || (n > 0 && (!ctr || !key || !out)))
        return fail(GFMC_EINVAL, "gfmc_philox_host: bad argument");
    for (int i = 0; i < n; ++i) {
        gfmc_u4 r = gfmc_philox(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3],
                                key[0], key[1]);
        out[4 * i] = r.x;
        out[4 * i + 1] = r.y;
        out[4 * i + 2] = r.z;
        out[4 * i + 3] = r.w;
    }
    return 0;
}

}  // extern "C"

#include "gfmc_train.inl"
#include "gfmc_batch.inl"
#include "gfmc_graph.inl"
#include "gfmc_estim.inl"
#include "gfmc_diag.inl"
#include "gfmc_mlp.inl"
#include "gfmc_mlp_train.inl"
#include "gfmc_offdiag.inl"
