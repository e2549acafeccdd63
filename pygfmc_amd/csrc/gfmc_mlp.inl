// gfmc_mlp.inl -- the two-layer MLP guide: log psi, rates and continuous-time
// propagation (included at the end of gfmc_hip.hip, one translation unit).
//
//   log psi(x) = b3 + sum_c w3_c s(b2_c + sum_k W2_ck h_k(x)),
//   h_k(x)     = bf16(s(z_k(x))),  z(x) = b1 + W1 x   (x = +-1),
//
// s the tabulated tanh.  A guide ratio changes every hidden unit, so each
// hop re-evaluates the network on every candidate move.  The propagation
// runs in bulk-synchronous iterations, one hop of every active walker each:
//   k_mlp_scan*   rows of each active walker (its configuration, then one row
//                 per slot with a non-zero bare rate), padded to 16-row
//                 blocks; an exclusive scan gives each walker its blocks;
//   k_mlp_gemm    one wavefront per 16-row block (one walker): the layer-1
//                 activations of its rows are built in registers straight in
//                 the MFMA A layout from z (exact integers at scale 2^16: W1
//                 and b1 lie on the layer-1 grid, BF16 values of magnitude
//                 2^-9 .. 64, so every z is a multiple of 2^-16 below 2^15)
//                 and the two W1 columns of the row's move, W2 is
//                 staged in LDS in 32-wide K chunks, v_mfma_f32_16x16x32_bf16
//                 accumulates the 16 x h2 block, the epilogue applies s, w3
//                 and a fixed-order row reduction;
//   k_mlp_select  one wavefront per walker: rates from the row values,
//                 e_L, the hold and the move exactly as the other
//                 continuous-time kernels (blocked order, Philox counter
//                 (hop, walker, step, tag)), then z and the candidate list of
//                 the new configuration.
// The host looks at the number of row blocks of the last scan once per 16
// hop iterations (one 4-byte copy) and stops when no walker is active; a
// step captured into a graph runs a fixed number of hop iterations instead
// (gfmc_mlp_set_hop_cap) and counts a step that left a walker unfinished
// (gfmc_mlp_unfinished).  Every value a configuration gets is a
// function of that configuration alone (exact z, fixed MFMA and reduction
// order for every row), so log psi(x') as a move candidate equals log psi(x')
// as the current configuration later.

// MLP_NT (2048): linear pieces of the tabulated tanh on [0, 8)
#define MLP_TSCALE 256.0f   // pieces per unit of |z|
#define MLP_GW 8            // wavefronts (32-row blocks) per GEMM workgroup
#define MLP_KC 32           // K chunk of W2 staged in LDS (one barrier per chunk)
#define MLP_RB 16           // rows per block (one MFMA row tile)
#define MLP_BSTR (MLP_KC + 8)
#define MLP_ZSCALE 65536.0  // z1 holds z x 2^16 (exact: layer-1 grid)
#define MLP_SAT (1u << 19)  // |z| x 2^16 at |z| = 8, where s saturates
// LDS of the GEMM kernels before W2: the FP32-route table, then the integer
// route's 2049 entries (padded to 2050)
#define MLP_TABS_BYTES ((2 * MLP_NT + 2) * 8)
// MLP_SW (4): walker wavefronts per start / select workgroup

typedef __attribute__((ext_vector_type(8))) __bf16 mlp_bf16x8;
typedef __attribute__((ext_vector_type(4))) float mlp_f32x4;

struct MlpWs {
    int32_t *z1;         // [cap][h1]   layer-1 pre-activations x 2^16 of the current config
    int32_t *cand;       // [cap][n_off] candidate slots, ascending
    int32_t *cnt;        // [cap]       rows: 1 + candidates
    int32_t *active;     // [cap]
    double *t, *lw, *eL; // [cap]
    uint32_t *it;        // [cap]
    int32_t *blk_off;    // [cap]       first 16-row block of the walker
    int32_t *blk_walker; // [cap * blocks per walker]
    int32_t *part;       // [cap / 1024 + 1] scan partials
    int32_t *total;      // [1]         row blocks of the iteration
    float *logpsi;       // [cap * blocks per walker * 16]
    long long *counters; // [5]         row blocks / hop iterations propagated (device sums),
                         //             captured steps left unfinished by the hop cap, the
                         //             largest hop-iteration count of a step, this step's
};

// rows of one training evaluation (gfmc_mlp_vmin_grad): the upstream
// coefficient of every row, its BF16 activations, its table indices (bit 15:
// saturated), its row record, the layer-2 upstream gradients, and per-wave
// partial sums of the w3 / b3 gradients
struct MlpTws {
    const float *coef;   // [rows]
    uint16_t *Hb;        // [rows][h1]
    uint16_t *Jx;        // [rows][h1]
    int2 *rinfo;         // [rows] (walker, si | sj << 16), 0xffff = none
    float *G2;           // [rows][h2]
    float *wpart;        // [waves][h2 + 1]
};

// the problem's view of the caller's propagation workspace (gfmc_mlp_set_work)
struct MlpWork {
    int64_t cap = 0;             // walkers the attached workspace holds (0: none attached)
    int bpw = 0;                 // blocks per walker at most
    int counting = 0;            // the propagation adds to ws.counters
    int hop_cap = 512;           // hop iterations of a step captured into a graph
    MlpWs ws{};
    int32_t *host_total = nullptr;   // pinned, allocated with the problem
};

// s(z) = sign(z) (T_j + f D_j): t = min(256 |z|, 2048 - 2^-13), j = floor(t),
// f = t - j (saturating at T_2047 + 0.99988 D_2047 ~ tanh(8))
#define MLP_TMAX 2047.9998779296875f

__device__ __forceinline__ float mlp_tanh(const float2 *tab, float z) {
    const float t = fminf(fabsf(z) * MLP_TSCALE, MLP_TMAX);
    const float2 e = tab[(int)t];
    return copysignf(__builtin_fmaf(__builtin_amdgcn_fractf(t), e.y, e.x), z);
}

__device__ __forceinline__ float bf16_val(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float mlp_bare(const DevProblem &P, const Tables<float> &tb,
                                          const signed char *x, int s) {
    if (s < P.n_hx) return tb.hx_half[s];
    const int e = s - P.n_hx;
    const int pk = P.edge_pack[e];
    return x[pk & 0xffff] != x[(pk >> 16) & 0xffff] ? tb.jpm[e] : tb.jpmpm[e];
}

// slots with a non-zero bare rate, ascending; returns their number
__device__ int mlp_candidates(const DevProblem &P, const Tables<float> &tb, const signed char *x,
                              int32_t *cand, int lane) {
    int base = 0;
    for (int c = 0; c < P.n_off; c += WAVE) {
        const int s = c + lane;
        const bool act = s < P.n_off && mlp_bare(P, tb, x, s) != 0.0f;
        const unsigned long long m = __ballot(act);
        if (act) cand[base + __popcll(m & ((1ull << lane) - 1ull))] = s;
        base += __popcll(m);
    }
    return base;
}

// z = b1 + W1 x from scratch (x 2^16, integer sums: exact in any order),
// eight consecutive hidden units per lane
__device__ void mlp_z_init(const DevProblem &P, const signed char *x, int32_t *z1, int lane) {
    for (int e0 = lane * 8; e0 < P.mlp_h1; e0 += WAVE * 8) {
        int32_t z[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) z[q] = (int32_t)(P.mlp_b1[e0 + q] * MLP_ZSCALE);
        for (int n = 0; n < P.N; ++n) {
            const int4 a = *(const int4 *)(P.mlp_W1I + (size_t)n * P.mlp_h1 + e0);
            const int4 b = *(const int4 *)(P.mlp_W1I + (size_t)n * P.mlp_h1 + e0 + 4);
            const int32_t wv[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            if (x[n] > 0) {
#pragma unroll
                for (int q = 0; q < 8; ++q) z[q] += wv[q];
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) z[q] -= wv[q];
            }
        }
        *(int4 *)(z1 + e0) = make_int4(z[0], z[1], z[2], z[3]);
        *(int4 *)(z1 + e0 + 4) = make_int4(z[4], z[5], z[6], z[7]);
    }
}

// z += the change of flipping site i (j >= 0 too) at the pre-flip x
__device__ void mlp_z_move(const DevProblem &P, const signed char *x, int32_t *z1, int i, int j,
                           int lane) {
    const int32_t ci = -2 * (int32_t)x[i], cj = j >= 0 ? -2 * (int32_t)x[j] : 0;
    for (int e = lane; e < P.mlp_h1; e += WAVE) {
        int32_t z = z1[e] + ci * P.mlp_W1I[(size_t)i * P.mlp_h1 + e];
        if (j >= 0) z += cj * P.mlp_W1I[(size_t)j * P.mlp_h1 + e];
        z1[e] = z;
    }
}

// start of a propagation / evaluation: spins to LDS, z from scratch, the
// candidate rows (stay_only: the configuration row alone), walker state
__global__ __launch_bounds__(MLP_SW * WAVE) void k_mlp_start(DevProblem P, Tables<float> tb,
                                                             int64_t nw, const uint32_t *spins,
                                                             const double *log_w, int stay_only,
                                                             MlpWs W) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    signed char *x = (signed char *)lds + (size_t)wave * P.n_pad;
    if (blockIdx.x == 0 && threadIdx.x == 0) W.counters[4] = 0;   // this step's hop iterations
    for (int64_t w = (int64_t)blockIdx.x * MLP_SW + wave; w < nw;
         w += (int64_t)gridDim.x * MLP_SW) {
        load_spins(P, spins + w * P.words, x, lane);
        wave_sync();
        mlp_z_init(P, x, W.z1 + (size_t)w * P.mlp_h1, lane);
        const int c = stay_only ? 0 : mlp_candidates(P, tb, x, W.cand + (size_t)w * P.n_off, lane);
        if (lane == 0) {
            W.cnt[w] = c + 1;
            W.active[w] = 1;
            W.t[w] = 0.0;
            W.lw[w] = log_w ? log_w[w] : 0.0;
            W.eL[w] = 0.0;
            W.it[w] = 0u;
        }
        wave_sync();
    }
}

__device__ __forceinline__ int mlp_block_excl_scan(int v, int *sh, int *total) {
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    int incl = v;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl += o;
    }
    if (lane == WAVE - 1) sh[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < (int)blockDim.x / WAVE; ++i) {
            const int t = sh[i];
            sh[i] = run;
            run += t;
        }
        sh[WAVE] = run;
    }
    __syncthreads();
    *total = sh[WAVE];
    return sh[wave] + incl - v;
}

__global__ __launch_bounds__(1024) void k_mlp_scan1(int64_t nw, MlpWs W) {
    __shared__ int sh[WAVE + 1];
    const int64_t w = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    const int nb = (w < nw && W.active[w]) ? (W.cnt[w] + MLP_RB - 1) / MLP_RB : 0;
    int tot;
    const int ex = mlp_block_excl_scan(nb, sh, &tot);
    if (w < nw) W.blk_off[w] = ex;
    if (threadIdx.x == 0) W.part[blockIdx.x] = tot;
}

// exclusive scan of the n workgroup totals in place; the sum to total[0]
__global__ __launch_bounds__(1024) void k_mlp_scan2(int n, MlpWs W) {
    __shared__ int sh[WAVE + 1];
    int carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int i = base + (int)threadIdx.x;
        const int v = i < n ? W.part[i] : 0;
        int tot;
        const int ex = mlp_block_excl_scan(v, sh, &tot);
        if (i < n) W.part[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) W.total[0] = carry;
}

// the propagation's counters: row blocks and (non-empty) hop iterations
__global__ void k_mlp_count(MlpWs W) {
    if (threadIdx.x == 0 && W.total[0] > 0) {
        W.counters[0] += W.total[0];
        W.counters[1] += 1;
        W.counters[4] += 1;
        if (W.counters[4] > W.counters[3]) W.counters[3] = W.counters[4];
    }
}

// after the capped hop iterations of a captured step: rows still to do
// mean walkers that did not finish their step
__global__ void k_mlp_capcheck(MlpWs W) {
    if (threadIdx.x == 0 && W.total[0] > 0) W.counters[2] += 1;
}

__global__ __launch_bounds__(1024) void k_mlp_scan3(int64_t nw, MlpWs W) {
    const int64_t w = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    if (w >= nw || !W.active[w]) return;
    const int off = W.blk_off[w] + W.part[blockIdx.x];
    W.blk_off[w] = off;
    const int nb = (W.cnt[w] + MLP_RB - 1) / MLP_RB;
    for (int q = 0; q < nb; ++q) W.blk_walker[off + q] = (int32_t)w;
}

// eight layer-1 activations of one row at hidden units k0 .. k0 + 7 from the
// prefetched z (FP64) and W1 columns (BF16) of the row's move (ci, cj: -2 x
// of the flipped sites, 0 when absent)
// c * w + z with 24-bit signed operands c and w, the sum wrapping (v_mad_i32_i24;
// written out because the compiler turns a chain of two into two multiplies
// and a three-operand add)
__device__ __forceinline__ int32_t mad_i24(int32_t c, int32_t w, int32_t z) {
    int32_t r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(c), "v"(w), "v"(z));
    return r;
}

struct MlpAIn {
    int4 z0, z1, i0, i1, j0, j1;   // z x 2^16 and the move's two W1 columns x 2^16
};

// rows past the walker's count load zeros: z = 0 gives the activation 0
__device__ __forceinline__ void mlp_a_load(MlpAIn &in, const int32_t *z1, const int32_t *w1i,
                                           const int32_t *w1j, int k0, bool rvalid) {
    if (rvalid) {
        in.z0 = *(const int4 *)(z1 + k0);
        in.z1 = *(const int4 *)(z1 + k0 + 4);
        in.i0 = *(const int4 *)(w1i + k0);
        in.i1 = *(const int4 *)(w1i + k0 + 4);
        in.j0 = *(const int4 *)(w1j + k0);
        in.j1 = *(const int4 *)(w1j + k0 + 4);
    } else {
        in.z0 = in.z1 = in.i0 = in.i1 = in.j0 = in.j1 = make_int4(0, 0, 0, 0);
    }
}

// The table of the integer route (staged in LDS after the FP32 one): entry
// j < 2048 is (T_j, D_j / 256), entry 2048 the saturated value
// (T_2047 + (1 - 2^-13) D_2047, 0).  With a = min(|z| 2^16, 2^19) the
// activation is T_j + f D_j / 256 at j = a >> 8, f = a & 255: t = 256 |z| =
// a / 256 exactly (z is a multiple of 2^-16 below 2^15, so the FP32 z of the
// FP32 route is exact wherever s does not saturate), floor(t) = j and
// fract(t) = f / 256, and fma(f, D / 256, T) rounds the same real number as
// fma(f / 256, D, T): bit for bit the value of mlp_tanh(z).
__device__ __forceinline__ void mlp_stage_tabi(const float2 *tab, float2 *tabi) {
    for (int i = threadIdx.x; i < MLP_NT; i += blockDim.x)
        tabi[i] = make_float2(tab[i].x, tab[i].y * 0.00390625f);
    if (threadIdx.x == 0) {
        const float2 e = tab[MLP_NT - 1];
        tabi[MLP_NT] = make_float2(__builtin_fmaf(MLP_TMAX - (float)(MLP_NT - 1), e.y, e.x), 0.0f);
    }
}

// the eight activations from the exact integer sums; the eight table reads
// are issued before any is used.  jout: the table index per activation (bit
// 15: saturated), for the slopes of the training rows
__device__ __forceinline__ mlp_bf16x8 mlp_a_build8(const int4 &za, const int4 &zb, const int4 &ia,
                                                   const int4 &ib, const int4 &ja, const int4 &jb,
                                                   int ci, int cj, const float2 *tabi,
                                                   uint4 *jout = nullptr) {
    const int32_t z0[8] = {za.x, za.y, za.z, za.w, zb.x, zb.y, zb.z, zb.w};
    const int32_t wi[8] = {ia.x, ia.y, ia.z, ia.w, ib.x, ib.y, ib.z, ib.w};
    const int32_t wj[8] = {ja.x, ja.y, ja.z, ja.w, jb.x, jb.y, jb.z, jb.w};
    int32_t z[8];
    uint32_t a[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        // |ci|, |cj| <= 2 and |w| < 2^23: two 24-bit multiply-adds (the sum
        // wraps for a padding row, which may hold any value; an exact z never does)
        z[q] = mad_i24(cj, wj[q], mad_i24(ci, wi[q], z0[q]));
        a[q] = min(z[q] < 0 ? 0u - (uint32_t)z[q] : (uint32_t)z[q], MLP_SAT);
    }
    float2 e[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) e[q] = tabi[__builtin_amdgcn_ubfe(a[q], 8, 12)];
    mlp_bf16x8 r;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float v = __builtin_fmaf((float)(a[q] & 255u), e[q].y, e[q].x);
        r[q] = (__bf16)copysignf(v, __int_as_float(z[q]));   // the sign bit of z
    }
    if (jout) {
        uint32_t jj[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) jj[q] = (a[q] >> 8) | (a[q] >= MLP_SAT ? 0x8000u : 0u);
        *jout = make_uint4(jj[0] | (jj[1] << 16), jj[2] | (jj[3] << 16), jj[4] | (jj[5] << 16),
                           jj[6] | (jj[7] << 16));
    }
    return r;
}

__device__ __forceinline__ mlp_bf16x8 mlp_a_build(const MlpAIn &in, int ci, int cj,
                                                  const float2 *tabi, uint4 *jout = nullptr) {
    return mlp_a_build8(in.z0, in.z1, in.i0, in.i1, in.j0, in.j1, ci, cj, tabi, jout);
}

// the move's two W1 columns (x 2^16) for one K step of the 32-row kernel
struct MlpWIn {
    int4 i0, i1, j0, j1;
};

__device__ __forceinline__ void mlp_w_load(MlpWIn &in, const int32_t *w1i, const int32_t *w1j,
                                           int k0) {
    in.i0 = *(const int4 *)(w1i + k0);
    in.i1 = *(const int4 *)(w1i + k0 + 4);
    in.j0 = *(const int4 *)(w1j + k0);
    in.j1 = *(const int4 *)(w1j + k0 + 4);
}

// the slope of the tabulated tanh at a table index (bit 15: saturated, 0)
__device__ __forceinline__ float mlp_slope(const float2 *tab, uint32_t j) {
    return (j & 0x8000u) ? 0.0f : MLP_TSCALE * tab[j & 0x7ff].y;
}

// the network on every row block; NC = h2 / 16 column blocks per wavefront
// (v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][8 (l / 16) + q], B[8 (l /
// 16) + q][l % 16], D register r = row 4 (l / 16) + r, column l % 16).
// MLP_GW wavefronts per workgroup share a double-buffered W2 chunk: chunk
// c + 1 is fetched into registers while chunk c feeds the MFMAs, then
// stored to the other buffer (one barrier per chunk); the A operand of K
// step s + 1 is loaded while step s is multiplied.
#ifndef MLP_WPE
#define MLP_WPE 2
#endif
template <int NC, bool TRAIN>
__global__ __launch_bounds__(MLP_GW * WAVE) __attribute__((amdgpu_waves_per_eu(MLP_WPE)))
void k_mlp_gemm(DevProblem P, const uint32_t *spins, MlpWs W, MlpTws T) {
    extern __shared__ __align__(16) unsigned char lds[];
    float2 *tab = (float2 *)lds;
    constexpr int H2 = NC * 16;
    constexpr int CH = H2 * (MLP_KC / 8);                 // uint4 per chunk
    constexpr int PER = (CH + MLP_GW * WAVE - 1) / (MLP_GW * WAVE);
    float2 *tabi = tab + MLP_NT;
    uint16_t *Bs0 = (uint16_t *)(lds + MLP_TABS_BYTES);
    uint16_t *Bs1 = Bs0 + H2 * MLP_BSTR;
    for (int i = threadIdx.x; i < MLP_NT; i += blockDim.x) tab[i] = P.mlp_tab[i];
    mlp_stage_tabi(P.mlp_tab, tabi);
    const int total = W.total[0];
    const int tiles = (total + MLP_GW - 1) / MLP_GW;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int H1 = P.mlp_h1, nchunks = H1 / MLP_KC;
    const int quad = lane >> 4, col = lane & 15;
    float acc3[NC], accb = 0.0f;   // TRAIN: this wave's w3 / b3 gradient partials
#pragma unroll
    for (int t = 0; t < NC; ++t) acc3[t] = 0.0f;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int b = tile * MLP_GW + wave;
        const bool bvalid = b < total;
        int64_t w = 0;
        int r0 = 0, cnt = 0;
        if (bvalid) {
            w = W.blk_walker[b];
            r0 = MLP_RB * (b - W.blk_off[w]);
            cnt = W.cnt[w];
        }
        const int row = r0 + col;
        const bool rvalid = bvalid && row < cnt;
        int si = -1, sj = -1;
        int ci = 0, cj = 0;
        if (rvalid && row > 0) {
            slot_sites(P, W.cand[(size_t)w * P.n_off + row - 1], &si, &sj);
            const uint32_t *sp = spins + w * P.words;
            ci = ((sp[si >> 5] >> (si & 31)) & 1u) ? -2 : 2;
            if (sj >= 0) cj = ((sp[sj >> 5] >> (sj & 31)) & 1u) ? -2 : 2;
        }
        const int32_t *z1 = W.z1 + (size_t)w * H1;
        const int32_t *w1i = P.mlp_W1I + (size_t)(si < 0 ? 0 : si) * H1;
        const int32_t *w1j = P.mlp_W1I + (size_t)(sj < 0 ? 0 : sj) * H1;
        const size_t grow = (size_t)b * MLP_RB + col;   // this lane's row (TRAIN stores)
        if (TRAIN && bvalid && quad == 0)
            T.rinfo[grow] = make_int2((int)w, (si & 0xffff) | ((sj & 0xffff) << 16));
        mlp_f32x4 acc[NC];
#pragma unroll
        for (int t = 0; t < NC; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[t][r] = 0.0f;
        // chunk 0 of W2 into buffer 0
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int idx = threadIdx.x + u * MLP_GW * WAVE;
            if (idx < CH) {
                const int j = idx / (MLP_KC / 8), q = idx % (MLP_KC / 8);
                *(uint4 *)(Bs0 + j * MLP_BSTR + q * 8) =
                    *(const uint4 *)(P.mlp_W2 + (size_t)j * H1 + q * 8);
            }
        }
        MlpAIn ain;
        mlp_a_load(ain, z1, w1i, w1j, 8 * quad, rvalid);
        __syncthreads();
        static_assert(PER <= 4, "W2 chunk prefetch holds four uint4 per thread");
        // the thread's chunk elements idx_u = threadIdx.x + u * blockDim.x
        const int TPB = MLP_GW * WAVE;
        const int jq = threadIdx.x / (MLP_KC / 8), qq = threadIdx.x % (MLP_KC / 8);
        const int jstep = TPB / (MLP_KC / 8);
        uint4 pre0 = make_uint4(0, 0, 0, 0), pre1 = pre0, pre2 = pre0, pre3 = pre0;
        for (int c = 0; c < nchunks; ++c) {
            const uint16_t *Bc = (c & 1) ? Bs1 : Bs0;
            uint16_t *Bn = (c & 1) ? Bs0 : Bs1;
            const bool more = c + 1 < nchunks;
            if (more) {
                const uint16_t *src = P.mlp_W2 + (size_t)jq * H1 + (c + 1) * MLP_KC + qq * 8;
                if ((int)threadIdx.x < CH) pre0 = *(const uint4 *)src;
                if (PER > 1 && (int)threadIdx.x + TPB < CH)
                    pre1 = *(const uint4 *)(src + (size_t)jstep * H1);
                if (PER > 2 && (int)threadIdx.x + 2 * TPB < CH)
                    pre2 = *(const uint4 *)(src + (size_t)2 * jstep * H1);
                if (PER > 3 && (int)threadIdx.x + 3 * TPB < CH)
                    pre3 = *(const uint4 *)(src + (size_t)3 * jstep * H1);
            }
#pragma unroll
            for (int ks = 0; ks < MLP_KC; ks += 32) {
                // build this K step's operand, then start the next step's loads
                uint4 jv;
                const mlp_bf16x8 a = mlp_a_build(ain, ci, cj, tabi, TRAIN ? &jv : nullptr);
                if (TRAIN && bvalid) {
                    const size_t off = grow * H1 + c * MLP_KC + ks + 8 * quad;
                    *(mlp_bf16x8 *)(T.Hb + off) = a;
                    *(uint4 *)(T.Jx + off) = jv;
                }
                const int knext = c * MLP_KC + ks + 32 + 8 * quad;
                if (knext < H1) mlp_a_load(ain, z1, w1i, w1j, knext, rvalid);
#pragma unroll
                for (int t = 0; t < NC; ++t) {
                    const mlp_bf16x8 bv =
                        *(const mlp_bf16x8 *)(Bc + (t * 16 + col) * MLP_BSTR + ks + 8 * quad);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
                }
            }
            if (more) {
                uint16_t *dst = Bn + jq * MLP_BSTR + qq * 8;
                if ((int)threadIdx.x < CH) *(uint4 *)dst = pre0;
                if (PER > 1 && (int)threadIdx.x + TPB < CH) *(uint4 *)(dst + jstep * MLP_BSTR) = pre1;
                if (PER > 2 && (int)threadIdx.x + 2 * TPB < CH)
                    *(uint4 *)(dst + 2 * jstep * MLP_BSTR) = pre2;
                if (PER > 3 && (int)threadIdx.x + 3 * TPB < CH)
                    *(uint4 *)(dst + 3 * jstep * MLP_BSTR) = pre3;
            }
            __syncthreads();
        }
        if constexpr (TRAIN) {
            // G2 = coef w3 s'(u) per (row, column); the w3 and b3 gradient
            // partials sum coef s(u) and coef over this wave's rows
            float cf[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                cf[r] = bvalid ? T.coef[(size_t)b * MLP_RB + 4 * quad + r] : 0.0f;
#pragma unroll
            for (int t = 0; t < NC; ++t) {
                const int j = t * 16 + col;
                const float bb = P.mlp_b2[j], ww = P.mlp_w3[j];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float u = acc[t][r] + bb;
                    const float tt = fminf(fabsf(u) * MLP_TSCALE, MLP_TMAX);
                    const uint32_t ji = (uint32_t)(int)tt | (tt == MLP_TMAX ? 0x8000u : 0u);
                    const float su = mlp_tanh(tab, u);
                    acc3[t] = acc3[t] + cf[r] * su;
                    if (bvalid)
                        T.G2[((size_t)b * MLP_RB + 4 * quad + r) * H2 + j] =
                            cf[r] * ww * mlp_slope(tab, ji);
                }
            }
            if (col == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) accb = accb + cf[r];
            continue;
        }
        // epilogue: per lane a sum over the column blocks t (column t * 16 +
        // col), then a butterfly over col
        float part[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) part[r] = 0.0f;
#pragma unroll
        for (int t = 0; t < NC; ++t) {
            const int j = t * 16 + col;
            const float bb = P.mlp_b2[j], ww = P.mlp_w3[j];
#pragma unroll
            for (int r = 0; r < 4; ++r) part[r] = part[r] + ww * mlp_tanh(tab, acc[t][r] + bb);
        }
#pragma unroll
        for (int d = 8; d >= 1; d >>= 1)
#pragma unroll
            for (int r = 0; r < 4; ++r) part[r] = part[r] + __shfl_xor(part[r], d, WAVE);
        if (bvalid && col == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * quad + r;
                if (r0 + i < cnt) W.logpsi[(size_t)b * MLP_RB + i] = part[r] + P.mlp_b3p[0];
            }
        }
    }
    if constexpr (TRAIN) {
        // quads hold the same columns for different rows: combine in order
#pragma unroll
        for (int t = 0; t < NC; ++t) {
            acc3[t] = acc3[t] + __shfl_xor(acc3[t], 16, WAVE);
            acc3[t] = acc3[t] + __shfl_xor(acc3[t], 32, WAVE);
        }
        accb = accb + __shfl_xor(accb, 16, WAVE);
        accb = accb + __shfl_xor(accb, 32, WAVE);
        float *wp = T.wpart + ((size_t)blockIdx.x * MLP_GW + wave) * (H2 + 1);
        if (quad == 0) {
#pragma unroll
            for (int t = 0; t < NC; ++t) wp[t * 16 + col] = acc3[t];
        }
        if (lane == 0) wp[H2] = accb;
    }
}

// The propagation's network rows in 32-row tiles: wavefront w of tile t
// takes the 16-row blocks 2 (t GW2 + w) and 2 (t GW2 + w) + 1 on
// v_mfma_f32_32x32x16_bf16 (lane l holds A[l % 32][8 (l / 32) + q] and
// B[8 (l / 32) + q][l % 32]; D register j is row 8 (j / 4) + 4 (l / 32) +
// j % 4, column l % 32), so every B fragment read from LDS feeds 32 rows
// (half the LDS traffic per row of k_mlp_gemm); a K step's B fragments are
// all in flight while its activations are built.  W2 is staged as in
// k_mlp_gemm (double-buffered 32-wide K chunks, one barrier per chunk).
// The accumulators start at b2, so u + b2 leaves the MFMAs.  Epilogue: per
// lane the sum over the column tiles t (column t 32 + l % 32) of w3 s(u + b2)
// (fma), then a fixed tree over the 32 lanes of a row set (a lane swap, then
// DPP): every row's value is the same sequence of operations wherever the
// row sits, so log psi(x') as a candidate row equals log psi(x') as a
// configuration row.
#define MLP_GW2 8
typedef __attribute__((ext_vector_type(16))) float mlp_f32x16;

// Operands of a row: z of its walker (ZLDS: the two blocks' z rows copied to
// LDS once per tile, read at each K step; otherwise read from global memory
// at each step) and the two W1 columns of its move, loaded two K steps ahead.
// Every lane loads: a padding row reads its walker's z and W1 row 0 (walker
// 0 for a block past the count), all in bounds, finite activations that reach
// only that row's MFMA outputs, which are never stored.
template <int NT, bool ZLDS>
__global__ __launch_bounds__(MLP_GW2 * WAVE) __attribute__((amdgpu_waves_per_eu(2)))
void k_mlp_gemm32(DevProblem P, const uint32_t *spins, MlpWs W) {
    // the tables in static LDS (addresses known when compiling: the table
    // reads take their offset as an immediate), W2's buffers in the dynamic part
    __shared__ float2 tab[MLP_NT];
    __shared__ float2 tabi[MLP_NT + 2];
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr int H2 = NT * 32;
    constexpr int CH = H2 * (MLP_KC / 8);                 // uint4 per chunk
    constexpr int TPB = MLP_GW2 * WAVE;
    constexpr int PER = (CH + TPB - 1) / TPB;
    static_assert(PER <= 4, "W2 chunk prefetch holds four uint4 per thread");
    uint16_t *Bs0 = (uint16_t *)lds;
    uint16_t *Bs1 = Bs0 + H2 * MLP_BSTR;
    for (int i = threadIdx.x; i < MLP_NT; i += blockDim.x) tab[i] = P.mlp_tab[i];
    mlp_stage_tabi(P.mlp_tab, tabi);
    const int total = W.total[0];
    const int pairs = (total + 1) >> 1;
    const int tiles = (pairs + MLP_GW2 - 1) / MLP_GW2;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int H1 = P.mlp_h1, nchunks = H1 / MLP_KC;
    int32_t *zs = (int32_t *)(Bs1 + H2 * MLP_BSTR) + (size_t)wave * 2 * H1;   // ZLDS: [2][H1]
    const int kh = lane >> 5, r32 = lane & 31;
    const int jq = threadIdx.x / (MLP_KC / 8), qq = threadIdx.x % (MLP_KC / 8);
    const int jstep = TPB / (MLP_KC / 8);
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int pb = tile * MLP_GW2 + wave;
        const int b = 2 * pb + (r32 >> 4);          // the block of this lane's A row
        const bool bvalid = b < total;
        int64_t w = 0;
        int r0 = 0, cnt = 0;
        if (bvalid) {
            w = W.blk_walker[b];
            r0 = MLP_RB * (b - W.blk_off[w]);
            cnt = W.cnt[w];
        }
        const int row = r0 + (r32 & 15);
        const bool rvalid = bvalid && row < cnt;
        int si = -1, sj = -1;
        int ci = 0, cj = 0;
        if (rvalid && row > 0) {
            slot_sites(P, W.cand[(size_t)w * P.n_off + row - 1], &si, &sj);
            const uint32_t *sp = spins + w * P.words;
            ci = ((sp[si >> 5] >> (si & 31)) & 1u) ? -2 : 2;
            if (sj >= 0) cj = ((sp[sj >> 5] >> (sj & 31)) & 1u) ? -2 : 2;
        }
        const int32_t *w1i = P.mlp_W1I + (size_t)(si < 0 ? 0 : si) * H1;
        const int32_t *w1j = P.mlp_W1I + (size_t)(sj < 0 ? 0 : sj) * H1;
        if constexpr (ZLDS) {
            // the z rows of the wave's two blocks (their walkers in lanes 0 and 16);
            // the previous tile's reads of zs came earlier in this wave's order
            const int wa = __shfl((int)w, 0, WAVE), wb = __shfl((int)w, 16, WAVE);
            for (int i = lane * 4; i < 2 * H1; i += WAVE * 4) {
                const int bh = i >= H1 ? 1 : 0;
                *(int4 *)(zs + i) =
                    *(const int4 *)(W.z1 + (size_t)(bh ? wb : wa) * H1 + (i - bh * H1));
            }
        }
        const int32_t *zg = W.z1 + (size_t)w * H1;          // !ZLDS: z in global memory
        const int32_t *zl = zs + (r32 >> 4) * H1;           // ZLDS: this lane's block's row
        // the accumulators start at the layer-2 bias of their column (t 32 + l % 32)
        mlp_f32x16 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const float b2c = P.mlp_b2[t * 32 + r32];
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[t][r] = b2c;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int idx = threadIdx.x + u * TPB;
            if (idx < CH) {
                const int j = idx / (MLP_KC / 8), q = idx % (MLP_KC / 8);
                *(uint4 *)(Bs0 + j * MLP_BSTR + q * 8) =
                    *(const uint4 *)(P.mlp_W2 + (size_t)j * H1 + q * 8);
            }
        }
        // the operand of step k: z of the row's walker and its move's W1 columns
        auto a_of = [&](int k, const MlpWIn &wv) {
            int4 za, zb;
            if constexpr (ZLDS) {
                za = *(const int4 *)(zl + k + 8 * kh);
                zb = *(const int4 *)(zl + k + 8 * kh + 4);
            } else {
                za = *(const int4 *)(zg + k + 8 * kh);
                zb = *(const int4 *)(zg + k + 8 * kh + 4);
            }
            return mlp_a_build8(za, zb, wv.i0, wv.i1, wv.j0, wv.j1, ci, cj, tabi);
        };
        // W1 columns of steps 0 and 1 (H1 >= 64), the operand of step 0, then
        // the stage of step 0 refilled with step 2's columns
        MlpWIn wa, wb;
        mlp_w_load(wa, w1i, w1j, 8 * kh);
        mlp_w_load(wb, w1i, w1j, 16 + 8 * kh);
        __syncthreads();
        mlp_bf16x8 a = a_of(0, wa);
        mlp_w_load(wa, w1i, w1j, 32 + 8 * kh);
        uint4 pre0 = make_uint4(0, 0, 0, 0), pre1 = pre0, pre2 = pre0, pre3 = pre0;
        // one K step, software-pipelined within the wavefront: the MFMAs of step
        // k interleaved with the build of step k + 1's operand (from the stage
        // wn, filled two steps before) and the refill of wn with step k + 3's
        // columns; branch-free (past the last step the indices are clamped to
        // the last one, in bounds, and what they produce is not used)
        auto kstep = [&](const uint16_t *Bc, int k, MlpWIn &wn) {
            mlp_bf16x8 bv[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t)
                bv[t] = *(const mlp_bf16x8 *)(Bc + (t * 32 + r32) * MLP_BSTR + (k % MLP_KC) +
                                              8 * kh);
            const mlp_bf16x8 an = a_of(min(k + 16, H1 - 16), wn);
            mlp_w_load(wn, w1i, w1j, min(k + 48, H1 - 16) + 8 * kh);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv[t], acc[t], 0, 0, 0);
            a = an;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);    // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 3, 0);    // LDS reads
                __builtin_amdgcn_sched_group_barrier(0x002, 12, 0);   // VALU
            }
        };
        for (int c = 0; c < nchunks; ++c) {
            const uint16_t *Bc = (c & 1) ? Bs1 : Bs0;
            uint16_t *Bn = (c & 1) ? Bs0 : Bs1;
            const bool more = c + 1 < nchunks;
            if (more) {
                const uint16_t *src = P.mlp_W2 + (size_t)jq * H1 + (c + 1) * MLP_KC + qq * 8;
                if ((int)threadIdx.x < CH) pre0 = *(const uint4 *)src;
                if (PER > 1 && (int)threadIdx.x + TPB < CH)
                    pre1 = *(const uint4 *)(src + (size_t)jstep * H1);
                if (PER > 2 && (int)threadIdx.x + 2 * TPB < CH)
                    pre2 = *(const uint4 *)(src + (size_t)2 * jstep * H1);
                if (PER > 3 && (int)threadIdx.x + 3 * TPB < CH)
                    pre3 = *(const uint4 *)(src + (size_t)3 * jstep * H1);
            }
            static_assert(MLP_KC == 32, "two K steps per W2 chunk");
            kstep(Bc, c * MLP_KC, wb);
            kstep(Bc, c * MLP_KC + 16, wa);
            if (more) {
                uint16_t *dst = Bn + jq * MLP_BSTR + qq * 8;
                if ((int)threadIdx.x < CH) *(uint4 *)dst = pre0;
                if (PER > 1 && (int)threadIdx.x + TPB < CH) *(uint4 *)(dst + jstep * MLP_BSTR) = pre1;
                if (PER > 2 && (int)threadIdx.x + 2 * TPB < CH)
                    *(uint4 *)(dst + 2 * jstep * MLP_BSTR) = pre2;
                if (PER > 3 && (int)threadIdx.x + 3 * TPB < CH)
                    *(uint4 *)(dst + 3 * jstep * MLP_BSTR) = pre3;
            }
            __syncthreads();
        }
        // epilogue: D register r holds row 8 (r / 4) + 4 kh + r % 4 of the tile,
        // column t 32 + l % 32; per lane the sum over t of w3 s(u) (one fma each)
        float part[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) part[r] = 0.0f;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const float ww = P.mlp_w3[t * 32 + r32];
#pragma unroll
            for (int r = 0; r < 16; ++r)
                part[r] = __builtin_fmaf(ww, mlp_tanh(tab, acc[t][r]), part[r]);
        }
        // the sum over the 32 lanes of each half: registers 2q and 2q + 1 paired
        // by v_permlane16_swap (row 0 of each half then holds lane l + lane
        // l + 16 of register 2q, row 1 of register 2q + 1), then a DPP tree over
        // the 16 lanes of the row (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
        // row_mirror): the same tree for every row of every tile, all in registers
        float red[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(part[2 * q]),
                                                             __float_as_uint(part[2 * q + 1]),
                                                             false, false);
            float v = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
            v = v + dpp_get<0xB1, 0xf>(v);
            v = v + dpp_get<0x4E, 0xf>(v);
            v = v + dpp_get<0x141, 0xf>(v);
            v = v + dpp_get<0x140, 0xf>(v);
            red[q] = v;
        }
        // the two blocks' row offsets and counts (held by lanes 0 and 16)
        const int r0a = __shfl(r0, 0, WAVE), cnta = __shfl(cnt, 0, WAVE);
        const int r0b = __shfl(r0, 16, WAVE), cntb = __shfl(cnt, 16, WAVE);
        if ((r32 & 15) == 0) {
            const int odd = r32 >> 4;    // lane 16 of the half holds the odd registers
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int r = 2 * q + odd;
                const int tr = 8 * (r >> 2) + 4 * kh + (r & 3);   // row of the tile
                const int hb = tr >> 4, i = tr & 15;
                const int bb = 2 * pb + hb;
                if (bb < total && (hb ? r0b : r0a) + i < (hb ? cntb : cnta))
                    W.logpsi[(size_t)bb * MLP_RB + i] = red[q] + P.mlp_b3p[0];
            }
        }
    }
}

// one hop of every active walker (the loop body of the continuous-time
// kernels) from the row values of this iteration
__global__ __launch_bounds__(MLP_SW * WAVE) void k_mlp_select(
    DevProblem P, Tables<float> tb, int64_t nw, uint32_t *spins, double *log_w, double *e_loc,
    double beta, double alw, uint32_t key0, uint32_t key1, uint32_t step, uint32_t tag,
    int32_t *evals, MlpWs W, const uint32_t *step_base) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    WalkerLds<float> L = carve<float>(P, lds + (size_t)wave * P.rbm_wave_bytes);
    step += step_base ? *step_base : 0u;
    for (int64_t w = (int64_t)blockIdx.x * MLP_SW + wave; w < nw;
         w += (int64_t)gridDim.x * MLP_SW) {
        if (!W.active[w]) continue;
        uint32_t *sp = spins + w * P.words;
        load_spins(P, sp, L.x, lane);
        for (int s = lane; s < P.K * WAVE; s += WAVE) L.w[s] = 0.0f;
        wave_sync();
        const size_t base = (size_t)W.blk_off[w] * MLP_RB;
        const int cnt = W.cnt[w];
        const float lp0 = W.logpsi[base];
        int32_t *cand = W.cand + (size_t)w * P.n_off;
        for (int c = lane; c < cnt - 1; c += WAVE) {
            const int s = cand[c];
            const float lr = W.logpsi[base + 1 + c] - lp0;
            L.w[wpos(s, P.K)] = -(mlp_bare(P, tb, L.x, s) * __expf(lr));
        }
        int cnt_anti = 0;
        if (P.has_jz) {
            for (int e = lane; e < P.n_edges; e += WAVE) {
                const int pk = P.edge_pack[e];
                const unsigned char a = L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff];
                if (P.flags) L.anti[e] = a;
                cnt_anti += a;
            }
        }
        wave_sync();
        const int n_anti = wave_sum_int(cnt_anti);
        const double diag = diag_value(P, L.x, L.anti, n_anti, lane);
        double t = W.t[w], lw = W.lw[w];
        uint32_t it = W.it[w];
        float lane_sum;
        const float incl = lane_scan(P, L, lane, &lane_sum);
        const float R = lane_bcast(incl, WAVE - 1);
        const double eL = diag - (double)R;
        const gfmc_u4 r = gfmc_philox(it, (uint32_t)w, step, tag, key0, key1);
        const double psi = gfmc_u53(r.z, r.w);
        int idx = -1;
        double tau_raw = INFINITY;
        if (hop_rate(R)) {
            float u = gfmc_u24(r.x) * R;
            const float cap = below(R);
            if (u > cap) u = cap;
            idx = sample_slot(P, L, incl, lane_sum, u, lane);
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
            mlp_z_move(P, L.x, W.z1 + (size_t)w * P.mlp_h1, i, j, lane);
            wave_sync();
            if (lane == 0) {
                L.x[i] = (signed char)-L.x[i];
                if (j >= 0) L.x[j] = (signed char)-L.x[j];
            }
            wave_sync();
            store_spins(P, sp, L.x, lane);
            const int c = mlp_candidates(P, tb, L.x, cand, lane);
            if (lane == 0) W.cnt[w] = c + 1;
        }
        const bool more = t < beta;
        if (lane == 0) {
            W.t[w] = t;
            W.lw[w] = lw;
            W.eL[w] = eL;
            W.it[w] = it;
            W.active[w] = more ? 1 : 0;
            if (!more) {
                log_w[w] = lw - alw * beta;
                e_loc[w] = eL;
                if (evals) evals[w] += (int32_t)it;
            }
        }
        wave_sync();
    }
}

// rates / log-differences / e_L / log psi of every walker from one
// evaluation's rows (gfmc_rates, gfmc_log_psi)
__global__ __launch_bounds__(MLP_SW * WAVE) void k_mlp_rates(DevProblem P, Tables<float> tb,
                                                             int64_t nw, const uint32_t *spins,
                                                             double *rates, double *e_loc,
                                                             double *log_diff, double *log_psi,
                                                             MlpWs W) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    WalkerLds<float> L = carve<float>(P, lds + (size_t)wave * P.rbm_wave_bytes);
    const int64_t M = P.n_off + 1;
    for (int64_t w = (int64_t)blockIdx.x * MLP_SW + wave; w < nw;
         w += (int64_t)gridDim.x * MLP_SW) {
        load_spins(P, spins + w * P.words, L.x, lane);
        for (int s = lane; s < P.K * WAVE; s += WAVE) L.w[s] = 0.0f;
        wave_sync();
        const size_t base = (size_t)W.blk_off[w] * MLP_RB;
        const int cnt = W.cnt[w];
        const float lp0 = W.logpsi[base];
        if (log_psi && lane == 0) log_psi[w] = (double)lp0;
        if (!(rates || e_loc || log_diff)) continue;
        const int32_t *cand = W.cand + (size_t)w * P.n_off;
        if (log_diff)
            for (int s = lane; s < P.n_off; s += WAVE) log_diff[w * M + 1 + s] = 0.0;
        wave_sync();
        for (int c = lane; c < cnt - 1; c += WAVE) {
            const int s = cand[c];
            const float lr = W.logpsi[base + 1 + c] - lp0;
            L.w[wpos(s, P.K)] = -(mlp_bare(P, tb, L.x, s) * __expf(lr));
            if (log_diff) log_diff[w * M + 1 + s] = (double)lr;
        }
        int cnt_anti = 0;
        if (P.has_jz) {
            for (int e = lane; e < P.n_edges; e += WAVE) {
                const int pk = P.edge_pack[e];
                const unsigned char a = L.x[pk & 0xffff] != L.x[(pk >> 16) & 0xffff];
                if (P.flags) L.anti[e] = a;
                cnt_anti += a;
            }
        }
        wave_sync();
        const int n_anti = wave_sum_int(cnt_anti);
        const double diag = diag_value(P, L.x, L.anti, n_anti, lane);
        float lane_sum;
        const float incl = lane_scan(P, L, lane, &lane_sum);
        const float R = lane_bcast(incl, WAVE - 1);
        if (rates) {
            if (lane == 0) rates[w * M] = diag;
            for (int s = lane; s < P.n_off; s += WAVE)
                rates[w * M + 1 + s] = -(double)L.w[wpos(s, P.K)];
        }
        if (log_diff && lane == 0) log_diff[w * M] = 0.0;
        if (e_loc && lane == 0) e_loc[w] = diag - (double)R;
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static MlpWork *new_mlp_work(int blocks_per_walker) {
    MlpWork *m = new MlpWork;
    m->bpw = blocks_per_walker;
    if (hipHostMalloc((void **)&m->host_total, 16, hipHostMallocDefault) != hipSuccess)
        m->host_total = nullptr;
    return m;
}

static void free_mlp_work(MlpWork *m) {
    if (!m) return;
    if (m->host_total) (void)hipHostFree(m->host_total);
    delete m;
}

// layout of the propagation workspace for nw walkers: 256-byte aligned
// pieces in the order of MlpWs; returns the total bytes (offsets into off)
static size_t mlp_work_layout(const DevProblem &P, int bpw, int64_t nw, size_t *off) {
    const size_t n = (size_t)nw, nb = n * (size_t)bpw;
    const size_t sizes[14] = {n * P.mlp_h1 * 4, n * (size_t)(P.n_off > 0 ? P.n_off : 1) * 4,
                              n * 4, n * 4, n * 8, n * 8, n * 8, n * 4, n * 4, nb * 4,
                              (n / 1024 + 2) * 4, 16, nb * MLP_RB * 4, 40};
    size_t o = 0;
    for (int i = 0; i < 14; ++i) {
        if (off) off[i] = o;
        o += (sizes[i] + 255) / 256 * 256;
    }
    return o;
}

static int mlp_reserve(const gfmc_problem *p, int64_t nw) {
    const MlpWork *m = p->mlp;
    if (m->cap >= nw) return 0;
    return fail(GFMC_EINVAL, "MLP guide: the attached workspace holds " + std::to_string(m->cap) +
                                 " walkers; attach gfmc_mlp_work_bytes(problem, " +
                                 std::to_string(nw) + ") bytes with gfmc_mlp_set_work");
}

static size_t mlp_walker_lds(const gfmc_problem *p) { return (size_t)MLP_SW * p->dp.rbm_wave_bytes; }

static size_t mlp_gemm_lds(const gfmc_problem *p) {
    return (size_t)MLP_TABS_BYTES + 2 * (size_t)p->dp.mlp_h2 * MLP_BSTR * 2;
}

static unsigned mlp_walker_grid(const gfmc_problem *p, int64_t nw) {
    const int64_t need = (nw + MLP_SW - 1) / MLP_SW;
    const int64_t cap = (int64_t)p->n_cu * 8;
    return (unsigned)(need < cap ? need : cap);
}

// the layer-2 GEMM over the row blocks the last scan counted (W.total, read
// on the device): a fixed persistent grid, so the host need not know the count
template <bool TRAIN>
static unsigned launch_mlp_gemm(const gfmc_problem *p, const uint32_t *spins, hipStream_t st,
                                const MlpTws &T = MlpTws{}) {
    const size_t lds = mlp_gemm_lds(p);
    const MlpWs &W = p->mlp->ws;
    if (!TRAIN) {
        // the propagation's and evaluations' rows: 32-row tiles (tables in static
        // LDS); the walkers' z rows in LDS as well when they fit beside W2's buffers
        const size_t zbytes = (size_t)MLP_GW2 * 2 * p->dp.mlp_h1 * 4;
        const bool zlds = lds + zbytes <= 160 * 1024;
        const size_t lds32 = lds - MLP_TABS_BYTES + (zlds ? zbytes : 0);
        const unsigned grid = (unsigned)(p->n_cu * 2);
        const dim3 g(grid), b(MLP_GW2 * WAVE);
#define MLP_G32(nt)                                                                          \
    do {                                                                                     \
        if (zlds)                                                                            \
            hipLaunchKernelGGL((k_mlp_gemm32<nt, true>), g, b, lds32, st, p->dp, spins, W);  \
        else                                                                                 \
            hipLaunchKernelGGL((k_mlp_gemm32<nt, false>), g, b, lds32, st, p->dp, spins, W); \
    } while (0)
        switch (p->dp.mlp_h2) {
        case 32: MLP_G32(1); break;
        case 64: MLP_G32(2); break;
        case 128: MLP_G32(4); break;
        default: MLP_G32(8); break;
        }
#undef MLP_G32
        return grid;
    }
    const unsigned grid = (unsigned)(p->n_cu * 4);
    const dim3 g(grid), b(MLP_GW * WAVE);
    switch (p->dp.mlp_h2) {
    case 32: hipLaunchKernelGGL((k_mlp_gemm<2, TRAIN>), g, b, lds, st, p->dp, spins, W, T); break;
    case 64: hipLaunchKernelGGL((k_mlp_gemm<4, TRAIN>), g, b, lds, st, p->dp, spins, W, T); break;
    case 128: hipLaunchKernelGGL((k_mlp_gemm<8, TRAIN>), g, b, lds, st, p->dp, spins, W, T); break;
    default: hipLaunchKernelGGL((k_mlp_gemm<16, TRAIN>), g, b, lds, st, p->dp, spins, W, T); break;
    }
    return grid;
}

// the row blocks of the active walkers, their count to W.total (device)
static void mlp_scan(const gfmc_problem *p, int64_t nw, hipStream_t st) {
    const MlpWs &W = p->mlp->ws;
    const unsigned nb = (unsigned)((nw + 1023) / 1024);
    hipLaunchKernelGGL(k_mlp_scan1, dim3(nb), dim3(1024), 0, st, nw, W);
    hipLaunchKernelGGL(k_mlp_scan2, dim3(1), dim3(1024), 0, st, (int)nb, W);
    hipLaunchKernelGGL(k_mlp_scan3, dim3(nb), dim3(1024), 0, st, nw, W);
}

// the count of the last scan on the host (one synchronisation)
static int mlp_read_total(const gfmc_problem *p, hipStream_t st, int *total) {
    if (!p->mlp->host_total) return fail(GFMC_ENOMEM, "MLP guide: no pinned host word");
    HIP_TRY(hipMemcpyAsync(p->mlp->host_total, p->mlp->ws.total, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *total = p->mlp->host_total[0];
    return 0;
}

// hop iterations enqueued between two looks at the row count
#define MLP_CHECK_EVERY 16

static int mlp_start(const gfmc_problem *p, int64_t nw, const uint32_t *spins,
                     const double *log_w, int stay_only, hipStream_t st) {
    int rc = mlp_reserve(p, nw);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_start, dim3(mlp_walker_grid(p, nw)), dim3(MLP_SW * WAVE),
                       (size_t)MLP_SW * p->dp.n_pad, st, p->dp, p->t32, nw, spins, log_w,
                       stay_only, p->mlp->ws);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int mlp_propagate(const gfmc_problem *p, int64_t nw, uint32_t *spins, double *log_w,
                         double *e_loc, double beta, double alw, uint32_t k0, uint32_t k1,
                         uint32_t step, uint32_t tag, int32_t *evals, hipStream_t st) {
    int rc = mlp_start(p, nw, spins, log_w, 0, st);
    if (rc) return rc;
    if (g_step_base) {
        // captured into a graph: a fixed number of hop iterations (walkers
        // that are done make them empty launches), then one more scan, whose
        // row count says whether a walker was left unfinished
        for (int iter = 0; iter < p->mlp->hop_cap; ++iter) {
            mlp_scan(p, nw, st);
            hipLaunchKernelGGL(k_mlp_count, dim3(1), dim3(WAVE), 0, st, p->mlp->ws);
            launch_mlp_gemm<false>(p, spins, st);
            hipLaunchKernelGGL(k_mlp_select, dim3(mlp_walker_grid(p, nw)),
                               dim3(MLP_SW * WAVE), mlp_walker_lds(p), st, p->dp, p->t32, nw,
                               spins, log_w, e_loc, beta, alw, k0, k1, step, tag, evals,
                               p->mlp->ws, g_step_base);
        }
        mlp_scan(p, nw, st);
        hipLaunchKernelGGL(k_mlp_capcheck, dim3(1), dim3(WAVE), 0, st, p->mlp->ws);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    // hop iterations of every active walker (scan -> GEMM -> select), the
    // GEMM and the select reading the row count on the device; after every
    // MLP_CHECK_EVERY iterations the host looks at the count of the last
    // scan (once all walkers are done the iterations are empty launches)
    for (int64_t iter = 0;; ++iter) {
        mlp_scan(p, nw, st);
        hipLaunchKernelGGL(k_mlp_count, dim3(1), dim3(WAVE), 0, st, p->mlp->ws);
        launch_mlp_gemm<false>(p, spins, st);
        hipLaunchKernelGGL(k_mlp_select, dim3(mlp_walker_grid(p, nw)), dim3(MLP_SW * WAVE),
                           mlp_walker_lds(p), st, p->dp, p->t32, nw, spins, log_w, e_loc, beta,
                           alw, k0, k1, step, tag, evals, p->mlp->ws, nullptr);
        HIP_TRY(hipGetLastError());
        if ((iter + 1) % MLP_CHECK_EVERY == 0) {
            int total = 0;
            rc = mlp_read_total(p, st, &total);
            if (rc) return rc;
            if (total == 0) break;
            if (iter >= 10000000)
                return fail(GFMC_EINVAL, "MLP guide: a walker did not finish its branching "
                                         "step in 10^7 hops (non-finite rates?)");
        }
    }
    return 0;
}

static int mlp_rates(const gfmc_problem *p, int64_t nw, const uint32_t *spins, double *rates,
                     double *e_loc, double *log_diff, double *log_psi, hipStream_t st) {
    if (g_step_base)
        return fail(GFMC_EUNSUP, "the MLP guide's evaluation cannot be captured into a graph");
    const int stay_only = !(rates || e_loc || log_diff);
    int rc = mlp_start(p, nw, spins, nullptr, stay_only, st);
    if (rc) return rc;
    mlp_scan(p, nw, st);
    launch_mlp_gemm<false>(p, spins, st);
    hipLaunchKernelGGL(k_mlp_rates, dim3(mlp_walker_grid(p, nw)), dim3(MLP_SW * WAVE),
                       mlp_walker_lds(p), st, p->dp, p->t32, nw, spins, rates, e_loc, log_diff,
                       log_psi, p->mlp->ws);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int gfmc_mlp_counters(const gfmc_problem *p, int64_t *out) {
    if (!p || !out) return fail(GFMC_EINVAL, "gfmc_mlp_counters: NULL argument");
    if (p->dp.guide != GFMC_GUIDE_MLP || !p->mlp)
        return fail(GFMC_EINVAL, "gfmc_mlp_counters: the problem's guide is not an MLP");
    if (p->mlp->cap == 0) {
        out[0] = out[1] = 0;
        return 0;
    }
    long long c[2];
    HIP_TRY(hipMemcpy(c, p->mlp->ws.counters, sizeof c, hipMemcpyDeviceToHost));
    out[0] = c[0];
    out[1] = c[1];
    return 0;
}

extern "C" int gfmc_mlp_set_hop_cap(gfmc_problem *p, int32_t hop_cap) {
    if (!p || p->dp.guide != GFMC_GUIDE_MLP || !p->mlp)
        return fail(GFMC_EINVAL, "gfmc_mlp_set_hop_cap: the problem's guide is not an MLP");
    if (hop_cap < 1) return fail(GFMC_EINVAL, "gfmc_mlp_set_hop_cap: hop_cap must be >= 1");
    p->mlp->hop_cap = hop_cap;
    return 0;
}

extern "C" int gfmc_mlp_unfinished(const gfmc_problem *p, int64_t *out) {
    if (!p || !out) return fail(GFMC_EINVAL, "gfmc_mlp_unfinished: NULL argument");
    if (p->dp.guide != GFMC_GUIDE_MLP || !p->mlp)
        return fail(GFMC_EINVAL, "gfmc_mlp_unfinished: the problem's guide is not an MLP");
    if (p->mlp->cap == 0) {
        out[0] = out[1] = 0;
        return 0;
    }
    long long c[2] = {0, 0};
    HIP_TRY(hipMemcpy(c, p->mlp->ws.counters + 2, sizeof c, hipMemcpyDeviceToHost));
    out[0] = c[0];
    out[1] = c[1];
    return 0;
}

extern "C" int64_t gfmc_mlp_work_bytes(const gfmc_problem *p, int64_t n_walkers) {
    if (!p || p->dp.guide != GFMC_GUIDE_MLP || !p->mlp || n_walkers < 1) return -1;
    if (n_walkers * p->mlp->bpw > 0x7fffffffLL / MLP_RB) return -1;
    return (int64_t)mlp_work_layout(p->dp, p->mlp->bpw, n_walkers, nullptr);
}

extern "C" int gfmc_mlp_set_work(gfmc_problem *p, void *work, int64_t bytes, int64_t n_walkers,
                                 void *stream) {
    if (!p || p->dp.guide != GFMC_GUIDE_MLP || !p->mlp)
        return fail(GFMC_EINVAL, "gfmc_mlp_set_work: the problem's guide is not an MLP");
    if (n_walkers < 1 || !work) return fail(GFMC_EINVAL, "gfmc_mlp_set_work: bad workspace");
    const int64_t need = gfmc_mlp_work_bytes(p, n_walkers);
    if (need < 0) return fail(GFMC_EUNSUP, "gfmc_mlp_set_work: too many walkers x candidate "
                                           "rows for one launch");
    if (bytes < need)
        return fail(GFMC_EINVAL, "gfmc_mlp_set_work: " + std::to_string(n_walkers) +
                                     " walkers need " + std::to_string(need) + " bytes");
    if (((uintptr_t)work & 255) != 0)
        return fail(GFMC_EINVAL, "gfmc_mlp_set_work: the workspace must be 256-byte aligned");
    size_t off[14];
    mlp_work_layout(p->dp, p->mlp->bpw, n_walkers, off);
    unsigned char *b = (unsigned char *)work;
    MlpWs &W = p->mlp->ws;
    W.z1 = (int32_t *)(b + off[0]);
    W.cand = (int32_t *)(b + off[1]);
    W.cnt = (int32_t *)(b + off[2]);
    W.active = (int32_t *)(b + off[3]);
    W.t = (double *)(b + off[4]);
    W.lw = (double *)(b + off[5]);
    W.eL = (double *)(b + off[6]);
    W.it = (uint32_t *)(b + off[7]);
    W.blk_off = (int32_t *)(b + off[8]);
    W.blk_walker = (int32_t *)(b + off[9]);
    W.part = (int32_t *)(b + off[10]);
    W.total = (int32_t *)(b + off[11]);
    W.logpsi = (float *)(b + off[12]);
    W.counters = (long long *)(b + off[13]);
    HIP_TRY(hipMemsetAsync(W.counters, 0, 40, (hipStream_t)stream));
    p->mlp->cap = n_walkers;
    return 0;
}
