// rbx_bf16.hip -- BF16-split form of the left dense sketch on the gfx950
// BF16 matrix cores (no RandBLAS counterpart; the FP32 analogue is lskge3,
// skge.hh:154-202).
//
//   B = alpha * op(bf16(submat(S))) * op(A) + beta * B
//
// Three kernels:
//   1. split:  op(A) (FP32, any strides) -> T bfloat16 planes A_1..A_T with
//      A = A_1 + ... + A_T + residual.  Every term is the round-to-nearest-even
//      of the previous remainder, and each remainder is exact in FP32.  The
//      planes are stored fragment-packed: for a 32-column block and a 16-deep
//      K step, the 64 lanes' MFMA B operands (8 bf16 each) are one contiguous
//      1 KB run, so every load instruction of the main kernel is a full
//      coalesced 1 KB read.  Zero padding to Kp (multiple of 32) x Np
//      (multiple of 256).
//   2. main:   one 256-thread workgroup per (64-row block, 1024-column block,
//      K-split).  Each 32-wide K-chunk of the operator is generated from its
//      Philox counters, rounded to bfloat16 and double-buffered in LDS; the
//      four waves each own 256 columns and keep a 64 x 256 FP32 accumulator
//      (2 x 8 blocks of v_mfma_f32_32x32x16_bf16, 256 AGPRs).  Each operator
//      entry is generated once per 1024 columns.  B operands rotate through
//      registers: the load of block j for the next K step is issued right
//      after block j's MFMAs of this step, so loads are in flight for a whole
//      step.  The K-split bounds each FP32 accumulation chain; partial tiles
//      go to a row-major scratch.
//   3. reduce: B = alpha * sum_s partial_s (in FP64, fixed order) + beta * B.
// WG-to-tile mapping is XCD-aware: workgroups that share a K-range are
// dispatched to the same XCD back to back, so the A planes they stream are
// served from that XCD's L2.
#include <vector>

#include "rbx_internal.hh"
#include "rbx_gemm_common.hh"
#include "rbx_fastgen.hh"


namespace rbx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

constexpr int BF_MT = 64, BF_NT = 1024, BF_KC = 32, BF_LDK = 40, BF_NTHR = 256;


// ---------------------------------------------------------------- split
// Element (k, j) of term t (of T) lives at
//   ((((j/32)*(Kp/16) + k/16)*T + t)*64 + j%32 + 32*((k/8)%2))*8 + k%8,
// i.e. for each 32-column block and 16-deep K step the T terms' 1 KB runs are
// adjacent: a kernel needs one pointer per column block, and the term is a
// constant 1 KB offset.
// One wave per (32-column block cb, 16-deep K step s).  Lane l handles column
// 32*cb + l%32 and k = 16*s + 8*(l/32) .. +7, and writes lane l's 16-byte slot
// of the block's 1 KB run of every term, so the wave's stores are T
// contiguous KB.
__global__ __launch_bounds__(256) void split_kernel(const float *A, int64_t rs, int64_t cs, int64_t K, int64_t N,
                                                    int64_t Kp, int64_t Np, int terms, uint16_t *planes) {
    const int64_t KS = Kp / 16;
    const int64_t n_waves = (Np / 32) * KS;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t wstride = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = wave0; w < n_waves; w += wstride) {
        const int64_t cb = w / KS, st = w - cb * KS;
        const int64_t j = cb * 32 + (lane & 31);
        const int64_t k0 = st * 16 + 8 * (lane >> 5);
        uint16_t out[3][8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int64_t k = k0 + e;
            float a = (k < K && j < N) ? A[k * rs + j * cs] : 0.0f;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                uint16_t h = bf16_rne(a);
                out[t][e] = h;
                a = __fsub_rn(a, bf16_to_f32(h));  // exact
            }
        }
        const int64_t dst = (cb * KS + st) * terms * 512 + lane * 8;
        for (int t = 0; t < terms; ++t) {
            u32x4 v;
            v.x = (uint32_t)out[t][0] | ((uint32_t)out[t][1] << 16);
            v.y = (uint32_t)out[t][2] | ((uint32_t)out[t][3] << 16);
            v.z = (uint32_t)out[t][4] | ((uint32_t)out[t][5] << 16);
            v.w = (uint32_t)out[t][6] | ((uint32_t)out[t][7] << 16);
            *(u32x4 *)(planes + dst + t * 512) = v;
        }
    }
}

// Unit-stride K (column-major A, or row-major A^T): one workgroup per
// (32-column block, 256-deep K range).  Each column's 256 consecutive values
// are one contiguous 1 KB read; the tile is transposed through LDS and then
// written as the block's 16 steps of fragment runs (the layout above).
constexpr int SPL_KT = 256, SPL_LD = 260;
template <bool VEC>
__global__ __launch_bounds__(256) void split_kunit_kernel(const float *A, int64_t cs, int64_t K, int64_t N,
                                                          int64_t Kp, int terms, uint16_t *planes,
                                                          int64_t kt_base) {
    __shared__ __attribute__((aligned(16))) float tile[32][SPL_LD];
    const int tid = threadIdx.x;
    const int64_t KS = Kp / 16;
    const int64_t cb = blockIdx.y;
    const int64_t kt0 = ((int64_t)blockIdx.x + kt_base) * SPL_KT;  // this launch's K tiles start at kt_base
    // load: a wave reads 1 KB runs of one column, k in [kt0, kt0 + 256); all
    // eight of a thread's loads are issued before the first LDS store.
    // VEC: 16-byte loads (A and the column stride are 4-float aligned)
    const int q = tid & 63, c0 = tid >> 6;
    const int64_t k = kt0 + 4 * q;
    float4 v[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const int64_t j = cb * 32 + c0 + 4 * p;
        const float *src = A + k + j * cs;
        if (VEC && j < N && k + 3 < K) {
            v[p] = *(const float4 *)src;
        } else {
            v[p].x = (j < N && k < K) ? src[0] : 0.0f;
            v[p].y = (j < N && k + 1 < K) ? src[1] : 0.0f;
            v[p].z = (j < N && k + 2 < K) ? src[2] : 0.0f;
            v[p].w = (j < N && k + 3 < K) ? src[3] : 0.0f;
        }
    }
#pragma unroll
    for (int p = 0; p < 8; ++p) *(float4 *)&tile[c0 + 4 * p][4 * q] = v[p];
    __syncthreads();
    // split: lane task (step s, lane l) of this K range, 4 per thread
    const int64_t steps = (KS - kt0 / 16) < 16 ? (KS - kt0 / 16) : 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int task = tid + 256 * i;
        const int st = task >> 6, l = task & 63;
        if (st >= steps) break;
        const float *src = &tile[l & 31][16 * st + 8 * (l >> 5)];
        // pairs rounded by the hardware RNE conversion (v_cvt_pk_bf16_f32,
        // the same bits as bf16_rne for every non-NaN input), each
        // remainder exact in FP32
        f32x2 a[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) a[p] = f32x2{src[2 * p], src[2 * p + 1]};
        u32x4 out[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(a[p], bf16x2));
                out[t][p] = h;
                a[p][0] = __fsub_rn(a[p][0], __uint_as_float(h << 16));
                a[p][1] = __fsub_rn(a[p][1], __uint_as_float(h & 0xffff0000u));
            }
        }
        const int64_t dst = (cb * KS + kt0 / 16 + st) * terms * 512 + l * 8;
        for (int t = 0; t < terms; ++t) *(u32x4 *)(planes + dst + t * 512) = out[t];
    }
}

struct Bf16Problem {
    int64_t M, N, K, Kp, Np;
    GenMap g;
    const uint16_t *planes;  // fragment-packed terms (see split_kernel)
    float *part;  // nS x M x N, row-major
    int64_t kw;   // K per split (multiple of BF_KC)
    int nMB, nNB, nS;  // row blocks, column blocks and K-splits of the launch
    int mb0;      // first row block of the launch
    int sp0;      // first K-split of the launch
};

// 32x32 output block, element r of lane l: row 8*(r/4) + 4*(l/32) + r%4, column l%32.
__device__ __forceinline__ int out_row32(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// FAST: the launched row blocks are interior (m0 + 64 <= M), K is a multiple
// of the 32-deep chunk (so every chunk is inside the operator), N is a
// multiple of 256 (every wave owns 8 valid column blocks), and the operator's
// rows run along K from a counter boundary (b_off % 4 == 0).  Then each thread
// owns two fixed Philox cells per chunk whose counters advance by 8 per chunk,
// and no bounds logic is compiled in.  Row blocks start at pb.mb0.
template <int TERMS, bool FAST>
__global__ __launch_bounds__(BF_NTHR, 1) void bf16_skge_kernel(Bf16Problem pb) {
    __shared__ __attribute__((aligned(16))) uint16_t Ss[2][BF_MT * BF_LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // XCD-aware tile order: consecutive tiles (same K-split) land on one XCD
    const int W = pb.nMB * pb.nNB * pb.nS;
    const int L = blockIdx.x;
    const int t = (W % 8 == 0) ? (L % 8) * (W / 8) + L / 8 : L;
    const int mb = t % pb.nMB;
    const int rest = t / pb.nMB;
    const int nb = rest % pb.nNB;
    const int sp = rest / pb.nNB;
    const int64_t m0 = (int64_t)(mb + pb.mb0) * BF_MT;
    const int64_t n0 = (int64_t)nb * BF_NT;
    const int64_t kbeg = (int64_t)sp * pb.kw;
    int64_t kend = kbeg + pb.kw;
    if (kend > pb.Kp) kend = pb.Kp;
    const int nchunks = (int)((kend - kbeg) / BF_KC);
    const GenMap g = pb.g;

    f32x16 acc[2][8];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int64_t wcol = n0 + wave * 256;        // first column of this wave
    const bool active = FAST || wcol < pb.Np;    // Np is a multiple of 256: all 8 blocks or none
    const int64_t KS = pb.Kp / 16;
    // lane's B operand for block j, K step s, term t:
    //   planes + (((wcol/32 + j)*KS + s)*TERMS + t)*512 + lane*8
    const uint16_t *bbase = pb.planes + (wcol >> 5) * KS * (TERMS * 512) + lane * 8;
    const int64_t bstep_blk = KS * (TERMS * 512);  // elements between column blocks
    const int64_t s_first = kbeg / 16, s_last = kend / 16 - 1;

    // FAST generator cells: thread owns cell (tid & 7) of operator rows
    // tid/8 and tid/8 + 32 of every chunk
    const int frow = tid >> 3, fcell = tid & 7;
    const uint64_t foff = FAST ? (uint64_t)((g.a_off + m0 + frow) * g.P + ((g.b_off + kbeg) >> 2) + fcell) : 0;
    const uint64_t foff2 = foff + (uint64_t)(32 * g.P);
    auto gen_chunk = [&](int c, uint16_t *lds) {
        if constexpr (FAST) {
            const bool gauss = g.gauss != 0;
#pragma unroll
            for (int it = 0; it < 2; ++it) {
                const float4 v = gen4(gauss, ctr_add(g.ctr, (it ? foff2 : foff) + 8 * (uint64_t)c), g.k0, g.k1);
                uint2 pk;
                pk.x = (uint32_t)entry_value<uint16_t>(v.x, gauss) |
                       ((uint32_t)entry_value<uint16_t>(v.y, gauss) << 16);
                pk.y = (uint32_t)entry_value<uint16_t>(v.z, gauss) |
                       ((uint32_t)entry_value<uint16_t>(v.w, gauss) << 16);
                *(uint2 *)(lds + (frow + 32 * it) * BF_LDK + fcell * 4) = pk;
            }
        } else {
            gen_tile<uint16_t, BF_MT, BF_KC, BF_NTHR>(g, m0, kbeg + (int64_t)c * BF_KC, pb.M, pb.K, lds, BF_LDK, 1,
                                                      tid);
        }
    };

    bf16x8 bq[8][TERMS];
    if (active) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int tt = 0; tt < TERMS; ++tt)
                bq[j][tt] = __builtin_bit_cast(
                    bf16x8, *(const u32x4 *)(bbase + j * bstep_blk + s_first * (TERMS * 512) + tt * 512));
    }

    if (nchunks > 0) gen_chunk(0, Ss[0]);
    __syncthreads();

    for (int c = 0; c < nchunks; ++c) {
        const int64_t kc = kbeg + (int64_t)c * BF_KC;
        if (c + 1 < nchunks) gen_chunk(c + 1, Ss[(c + 1) & 1]);
        const uint16_t *S = Ss[c & 1];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int koff = ks * 16 + 8 * (lane >> 5);
            const bf16x8 a0 = __builtin_bit_cast(bf16x8, *(const u32x4 *)(S + (lane & 31) * BF_LDK + koff));
            const bf16x8 a1 = __builtin_bit_cast(bf16x8, *(const u32x4 *)(S + (32 + (lane & 31)) * BF_LDK + koff));
            int64_t s_next = kc / 16 + ks + 1;
            if (s_next > s_last) s_next = s_last;  // redundant in-range reload at the end
            if (active) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
#pragma unroll
                    for (int tt = 0; tt < TERMS; ++tt) {
                        acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bq[j][tt], acc[0][j], 0, 0, 0);
                        acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bq[j][tt], acc[1][j], 0, 0, 0);
                    }
#pragma unroll
                    for (int tt = 0; tt < TERMS; ++tt)
                        bq[j][tt] = __builtin_bit_cast(
                            bf16x8, *(const u32x4 *)(bbase + j * bstep_blk + s_next * (TERMS * 512) + tt * 512));
                    // keep the reload of block j here, a whole step ahead of
                    // its use: without the barrier the scheduler sinks it
                    // next to the use and every block waits out the latency
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        __syncthreads();
    }

    // partial tile -> scratch (row-major M x N per split; lanes walk columns)
    if (!active) return;
    float *P = pb.part + (int64_t)sp * pb.M * pb.N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = m0 + i * 32 + out_row32(lane, r);
                const int64_t gj = wcol + j * 32 + (lane & 31);
                if (FAST || (gi < pb.M && gj < pb.N)) P[gi * pb.N + gj] = acc[i][j][r];
            }
}

// Interior tiles, two waves per SIMD.  A 512-thread workgroup owns a
// (32 RB) x (256 CB) output tile; each of its 8 waves owns all 32 RB rows and
// 32 CB columns of it (RB x CB blocks of 32 x 32, 16 RB CB accumulator
// registers, so two waves fit on a SIMD).  The B operands a wave streams per
// 16-deep step are CB x TERMS runs of 1 KB, so the global-load volume per
// MFMA falls as RB grows; the operator rows are generated once per
// workgroup, so generation per output column grows as CB shrinks.  Every
// thread generates RB/2 Philox cells of the next 32-deep chunk per chunk;
// waves 0-3 generate before their MFMAs and waves 4-7 after them, so the two
// waves sharing a SIMD (w and w + 4) are in opposite phases.  The same shape
// conditions as the FAST form of bf16_skge_kernel apply, with the tile
// (32 RB) x (256 CB).
constexpr int BF_FTHR = 512;
#ifndef RBX_BF16_TALL_RB
#define RBX_BF16_TALL_RB 2
#endif
constexpr int BF_RB = RBX_BF16_TALL_RB, BF_CB = 8 / BF_RB;  // blocks per wave: RB x CB = 8
// Wave priority while a wave issues its MFMAs (0: unchanged).
#ifndef RBX_BF16_PRIO
#define RBX_BF16_PRIO 1
#endif

// B-operand buffering depth of the pair kernel, in 16-deep K steps.
#ifndef RBX_BF16_BD
#define RBX_BF16_BD 1
#endif
// K depth of the operator chunks the pair kernel generates per barrier.
#ifndef RBX_BF16_KC
#define RBX_BF16_KC 32
#endif
constexpr int PKC = RBX_BF16_KC, PLDK = PKC + 8, PCPR = PKC / 4;  // PCPR: Philox cells per operator row

template <int TERMS, bool GAUSS, int RB, int CB>
__global__ __launch_bounds__(BF_FTHR, 1) void bf16_skge_pair_kernel(Bf16Problem pb) {
    constexpr int MT = 32 * RB, NT = 256 * CB;
    constexpr int RPP = BF_FTHR / PCPR, NCELL = MT / RPP;  // rows per generation pass, cells per thread
    __shared__ __attribute__((aligned(16))) uint16_t Ss[2][MT * PLDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // XCD-aware tile order: consecutive tiles (same K-split) land on one XCD
    const int W = pb.nMB * pb.nNB * pb.nS;
    const int L = blockIdx.x;
    const int t = (W % 8 == 0) ? (L % 8) * (W / 8) + L / 8 : L;
    const int mb = t % pb.nMB;
    const int rest = t / pb.nMB;
    const int nb = rest % pb.nNB;
    const int sp = rest / pb.nNB + pb.sp0;
    const int64_t m0 = (int64_t)(mb + pb.mb0) * MT;
    const int64_t n0 = (int64_t)nb * NT;
    const int64_t kbeg = (int64_t)sp * pb.kw;
    int64_t kend = kbeg + pb.kw;
    if (kend > pb.Kp) kend = pb.Kp;
    const int nchunks = (int)((kend - kbeg) / PKC);
    const GenMap g = pb.g;

    f32x16 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int64_t wcol = n0 + wave * (32 * CB);
    const int64_t KS = pb.Kp / 16;
    // lane's B operand for block j, K step s, term t:
    //   planes + (((wcol/32 + j)*KS + s)*TERMS + t)*512 + lane*8
    const uint16_t *bbase = pb.planes + (wcol >> 5) * KS * (TERMS * 512) + lane * 8;
    const int64_t bstep_blk = KS * (TERMS * 512);
    const int64_t s_first = kbeg / 16, s_last = kend / 16 - 1;

    // generator: thread owns cell (tid % PCPR) of operator rows tid / PCPR + RPP i
    const int frow = tid / PCPR, fcell = tid % PCPR;
    const uint64_t foff = (uint64_t)((g.a_off + m0 + frow) * g.P + ((g.b_off + kbeg) >> 2) + fcell);
    auto gen_chunk = [&](int c, uint16_t *lds) {
#pragma unroll
        for (int i = 0; i < NCELL; ++i) {
            *(uint2 *)(lds + (frow + RPP * i) * PLDK + fcell * 4) = gen4_bf16<GAUSS>(
                ctr_add(g.ctr, foff + (uint64_t)(RPP * i) * g.P + (uint64_t)PCPR * (uint64_t)c), g.k0, g.k1);
            __builtin_amdgcn_sched_barrier(0);
        }
    };

    // first B operands, loaded term-major: the order the loop reloads them
    // in, so the wait before each MFMA group covers only that group's load
    // B operands, BD steps deep: step s uses buffer s % BD (BD divides the
    // chunk's two steps), reloaded right after its MFMAs with step s + BD
    constexpr int BD = RBX_BF16_BD;
    static_assert(BD == 1 || BD == 2, "B operand depth");
    bf16x8 bq[BD][CB][TERMS];
#pragma unroll
    for (int bd = 0; bd < BD; ++bd)
#pragma unroll
        for (int tt = 0; tt < TERMS; ++tt)
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                int64_t s_ld = s_first + bd;
                if (s_ld > s_last) s_ld = s_last;
                bq[bd][j][tt] = __builtin_bit_cast(
                    bf16x8, *(const u32x4 *)(bbase + j * bstep_blk + s_ld * (TERMS * 512) + tt * 512));
            }

    // one 16-deep K step: block j's B operands for step s + BD are reloaded
    // right after its MFMAs, BD steps ahead of their use
    auto load_a = [&](const uint16_t *S, int ks, bf16x8 *a) {
        const int koff = ks * 16 + 8 * (lane >> 5);
#pragma unroll
        for (int i = 0; i < RB; ++i)
            a[i] = __builtin_bit_cast(bf16x8, *(const u32x4 *)(S + (32 * i + (lane & 31)) * PLDK + koff));
    };
    auto step = [&](const bf16x8 *a, int ks, int64_t s_next) {
        const int bb = ks % BD;
#if RBX_BF16_PRIO
        __builtin_amdgcn_s_setprio(RBX_BF16_PRIO);
#endif
        // term-outer: consecutive MFMAs on one accumulator are RB x CB apart,
        // and each accumulator still adds its terms in order hi, lo, lo2
#pragma unroll
        for (int tt = 0; tt < TERMS; ++tt) {
#pragma unroll
            for (int j = 0; j < CB; ++j) {
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], bq[bb][j][tt], acc[i][j], 0, 0, 0);
                bq[bb][j][tt] = __builtin_bit_cast(
                    bf16x8, *(const u32x4 *)(bbase + j * bstep_blk + s_next * (TERMS * 512) + tt * 512));
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#if RBX_BF16_PRIO
        __builtin_amdgcn_s_setprio(0);
#endif
    };

    const bool late = (wave & 4) != 0;
    gen_chunk(0, Ss[0]);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const uint16_t *S = Ss[c & 1];
        uint16_t *Sn = Ss[(c + 1) & 1];
        const int64_t s0 = s_first + (PKC / 16) * c;
        const bool more = c + 1 < nchunks;
        if (more && !late) gen_chunk(c + 1, Sn);
        // both steps' operator fragments are read from LDS up front, so the
        // second step's reads overlap the first step's MFMAs
        bf16x8 a[PKC / 16][RB];
#pragma unroll
        for (int ks = 0; ks < PKC / 16; ++ks) load_a(S, ks, a[ks]);
#pragma unroll
        for (int ks = 0; ks < PKC / 16; ++ks) step(a[ks], ks, s0 + ks + BD > s_last ? s_last : s0 + ks + BD);
        if (more && late) gen_chunk(c + 1, Sn);
        __syncthreads();
    }

    // partial tile -> scratch (row-major M x N per split; lanes walk columns)
    float *P = pb.part + (int64_t)sp * pb.M * pb.N;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = m0 + i * 32 + out_row32(lane, r);
                const int64_t gj = wcol + j * 32 + (lane & 31);
                P[gi * pb.N + gj] = acc[i][j][r];
            }
}

// B[i,j] = alpha * sum_s part[s][i*N + j] + beta * B[i*brs + j*bcs], sums in
// FP64 in split order.  RED_R x RED_C tiles (8 x 128): each row of a tile is
// one 512-byte run of every partial (FP32), long enough for full HBM bursts;
// a column-major B is written along i after a transpose through LDS.  A 1-D
// grid of tiles, row-tile major.  With pre != nullptr the sums start from
// pre[i*N + j], the FP64 sum of splits 0 .. s0 - 1 (bf16_presum_kernel), and
// continue with split s0: the same additions in the same order.
constexpr int RED_R = 8, RED_C = 128;
template <typename PT, typename OT>
__global__ __launch_bounds__(256) void bf16_reduce_kernel(const PT *part, int nS, int64_t M, int64_t N,
                                                          double alpha, double beta, OT *B, int64_t brs,
                                                          int64_t bcs, const double *pre = nullptr,
                                                          int s0 = 0) {
    __shared__ double tile[RED_R][RED_C + 1];
    const int64_t ntc = (N + RED_C - 1) / RED_C;
    const int64_t i0 = (int64_t)(blockIdx.x / ntc) * RED_R, j0 = (int64_t)(blockIdx.x % ntc) * RED_C;
    const int64_t total = M * N;
    {
        // thread: row r, columns c4 .. c4 + 3, 16-byte loads when whole and
        // aligned; eight splits' loads are issued together and every sum stays
        // in split order
        const int r = threadIdx.x / (RED_C / 4), c4 = 4 * (threadIdx.x % (RED_C / 4));
        const int64_t i = i0 + r, j = j0 + c4;
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        if (i < M) {
            if (pre != nullptr)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (j + e < N) v[e] = pre[i * N + j + e];
            const PT *p = part + i * N + j;
            if (j + 3 < N && N % 4 == 0) {
                typedef PT pt4 __attribute__((ext_vector_type(4)));
                int s = s0;
                for (; s + 8 <= nS; s += 8) {
                    pt4 x[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) x[u] = *(const pt4 *)(p + (int64_t)(s + u) * total);
#pragma unroll
                    for (int u = 0; u < 8; ++u)
#pragma unroll
                        for (int e = 0; e < 4; ++e) v[e] += (double)x[u][e];
                }
                for (; s < nS; ++s) {
                    const pt4 x = *(const pt4 *)(p + (int64_t)s * total);
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] += (double)x[e];
                }
            } else {
                for (int e = 0; e < 4; ++e)
                    if (j + e < N)
                        for (int s = s0; s < nS; ++s) v[e] += (double)p[(int64_t)s * total + e];
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) tile[r][c4 + e] = v[e];
    }
    __syncthreads();
    const bool cm = brs == 1;
#pragma unroll
    for (int q = 0; q < RED_R * RED_C / 256; ++q) {
        // column-major: consecutive threads walk i; row-major: they walk j
        const int e = threadIdx.x + 256 * q;
        const int ti = cm ? e % RED_R : e / RED_C, tj = cm ? e / RED_R : e % RED_C;
        const int64_t i = i0 + ti, j = j0 + tj;
        if (i < M && j < N) {
            const double v = tile[ti][tj];
            OT *b = B + i * brs + j * bcs;
            double out = alpha * v;
            if (beta != 0.0) out += beta * (double)(*b);
            *b = (OT)out;
        }
    }
}

// pre[e] = sum of part[s][e] over s = 0 .. ns - 1 in split order, FP64: the
// first splits of bf16_reduce_kernel's sums, run on a helper stream while the
// last K range is still being multiplied.  Lean (four elements and four
// partial loads in flight per thread), so its waves fit in the registers the
// main kernel leaves free on each SIMD.
typedef float pf32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void bf16_presum_kernel(const float *part, int ns, int64_t MN, double *pre) {
    const int64_t n4 = MN / 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += (int64_t)gridDim.x * blockDim.x) {
        const pf32x4 *p = (const pf32x4 *)part + q;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
        int s = 0;
        for (; s + 4 <= ns; s += 4) {
            pf32x4 x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = __builtin_nontemporal_load(p + (int64_t)(s + u) * n4);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v0 += (double)x[u][0];
                v1 += (double)x[u][1];
                v2 += (double)x[u][2];
                v3 += (double)x[u][3];
            }
        }
        for (; s < ns; ++s) {
            const pf32x4 x = __builtin_nontemporal_load(p + (int64_t)s * n4);
            v0 += (double)x[0];
            v1 += (double)x[1];
            v2 += (double)x[2];
            v3 += (double)x[3];
        }
        double *o = pre + 4 * q;
        o[0] = v0;
        o[1] = v1;
        o[2] = v2;
        o[3] = v3;
    }
}

// 1-D grid of bf16_reduce_kernel over an M x N output
static dim3 reduce_grid(int64_t M, int64_t N) {
    const int64_t nt = ((M + RED_R - 1) / RED_R) * ((N + RED_C - 1) / RED_C);
    RBX_REQUIRE(nt < (1ll << 31));
    return dim3((unsigned)nt);
}

template <int CB>
static void launch_pair(int n_terms, bool gs, dim3 grid, const Bf16Problem &q, hipStream_t stream) {
    if (n_terms == 1 && gs) hipLaunchKernelGGL((bf16_skge_pair_kernel<1, true, BF_RB, CB>), grid, dim3(BF_FTHR), 0, stream, q);
    else if (n_terms == 1) hipLaunchKernelGGL((bf16_skge_pair_kernel<1, false, BF_RB, CB>), grid, dim3(BF_FTHR), 0, stream, q);
    else if (n_terms == 2 && gs) hipLaunchKernelGGL((bf16_skge_pair_kernel<2, true, BF_RB, CB>), grid, dim3(BF_FTHR), 0, stream, q);
    else if (n_terms == 2) hipLaunchKernelGGL((bf16_skge_pair_kernel<2, false, BF_RB, CB>), grid, dim3(BF_FTHR), 0, stream, q);
    else if (gs) hipLaunchKernelGGL((bf16_skge_pair_kernel<3, true, BF_RB, CB>), grid, dim3(BF_FTHR), 0, stream, q);
    else hipLaunchKernelGGL((bf16_skge_pair_kernel<3, false, BF_RB, CB>), grid, dim3(BF_FTHR), 0, stream, q);
    RBX_HIP(hipGetLastError());
}

static void bf16_split_sketch(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                              const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const float *A, int64_t lda,
                              float beta, float *B, int64_t ldb, int n_terms, hipStream_t stream) {
    RBX_REQUIRE(S != nullptr);
    check_layout(layout);
    check_op(opS);
    check_op(opA);
    RBX_REQUIRE(n_terms >= 1 && n_terms <= 3);
    RBX_REQUIRE(d >= 0 && n >= 0 && m >= 0);
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
    if (S->buff != nullptr)
        RBX_FAIL(RBX_UNSUPPORTED, "bf16split: only generated operators (S.buff == NULL) are supported");
    DenseGeom geo = dense_geom(S->dist);
    safe_mul(geo.dim_minor, geo.P);
    int64_t rows_S = opS == 'N' ? d : m, cols_S = opS == 'N' ? m : d;
    RBX_REQUIRE(S->dist.n_rows >= rows_S + ro_s);
    RBX_REQUIRE(S->dist.n_cols >= cols_S + co_s);
    int64_t rows_A = opA == 'N' ? m : n, cols_A = opA == 'N' ? n : m;
    if (layout == 'C') {
        RBX_REQUIRE(lda >= rows_A);
        RBX_REQUIRE(ldb >= d);
    } else {
        RBX_REQUIRE(lda >= cols_A);
        RBX_REQUIRE(ldb >= n);
    }
    const int64_t M = d, N = n, K = m;
    const int64_t brs = layout == 'C' ? 1 : ldb, bcs = layout == 'C' ? ldb : 1;
    if (M == 0 || N == 0) return;
    if (alpha == 0.0f || K == 0) {
        scale_output<float>(B, M, N, brs, bcs, beta, stream);
        return;
    }
    // op(A)[k, j] strides
    int64_t ars, acs;
    if (layout == 'C') {
        ars = opA == 'N' ? 1 : lda;
        acs = opA == 'N' ? lda : 1;
    } else {
        ars = opA == 'N' ? lda : 1;
        acs = opA == 'N' ? 1 : lda;
    }
    Bf16Problem pb{};
    pb.M = M;
    pb.N = N;
    pb.K = K;
    pb.Kp = (K + BF_KC - 1) / BF_KC * BF_KC;
    pb.Np = (N + 255) / 256 * 256;
    pb.g = gen_map(*S, opS, ro_s, co_s, false);
    pb.nMB = (int)((M + BF_MT - 1) / BF_MT);
    pb.nNB = (int)((N + BF_NT - 1) / BF_NT);
    // K per split: bounded FP32 accumulation chains, and enough workgroups
    int64_t kw = 4096;
    int64_t tiles = (int64_t)pb.nMB * pb.nNB;
    int64_t want_splits = (4 * (int64_t)num_cus() + tiles - 1) / tiles;
    int64_t kw_fill = (pb.Kp + want_splits - 1) / want_splits;
    if (kw_fill < kw) kw = kw_fill;
    kw = (kw + PKC - 1) / PKC * PKC;  // a whole number of the pair kernel's chunks (and of BF_KC)
    if (kw < PKC) kw = PKC;
    pb.kw = kw;
    pb.nS = (int)((pb.Kp + kw - 1) / kw);
    RBX_REQUIRE((int64_t)pb.nMB * pb.nNB * pb.nS < (1ll << 31));


    // interior row blocks on the two-wave kernel when the shape allows it, the
    // rest (or everything) on the bounds-checked kernel; the widest tile (256
    // CB columns, CB = BF_CB, BF_CB / 2, BF_CB / 4) that divides N: narrow
    // sketches keep the two-wave kernel
    constexpr int TMT = 32 * BF_RB;
    int cb = 0;
    for (int c = BF_CB; c >= 1; c /= 2)
        if (N % (256 * c) == 0) {
            cb = c;
            break;
        }
    const bool fast_ok = K % PKC == 0 && cb > 0 && !pb.g.b_along_rows && (pb.g.b_off & 3) == 0;
    const int nfast = fast_ok ? (int)(M / TMT) : 0;   // row blocks of TMT rows
    const int nMB_all = pb.nMB;                        // row blocks of BF_MT rows
    const int nfast64 = nfast * (TMT / BF_MT);
    const bool vec = acs % 4 == 0 && ((uintptr_t)A & 15) == 0;
    const int64_t kt_total = (pb.Kp + SPL_KT - 1) / SPL_KT;
    RBX_REQUIRE(kt_total < (1ll << 31) && pb.Np / 32 <= 65535);
    // unit-stride K: coalesced column reads, transposed through LDS; K tiles
    // [kt_first, kt_first + kt_count) on stream s
    auto split_kunit = [&](int64_t kt_first, int64_t kt_count, hipStream_t s) {
        ProfScope prof("bf16_split_kernel", s);
        const dim3 grid((unsigned)kt_count, (unsigned)(pb.Np / 32));
        if (vec)
            hipLaunchKernelGGL(split_kunit_kernel<true>, grid, dim3(256), 0, s, A, acs, K, N, pb.Kp, n_terms,
                               (uint16_t *)pb.planes, kt_first);
        else
            hipLaunchKernelGGL(split_kunit_kernel<false>, grid, dim3(256), 0, s, A, acs, K, N, pb.Kp, n_terms,
                               (uint16_t *)pb.planes, kt_first);
        RBX_HIP(hipGetLastError());
    };
    // the two-wave kernel over K-splits [sp_first, sp_first + sp_count)
    auto main_fast = [&](int sp_first, int sp_count) {
        ProfScope prof("bf16_skge_kernel", stream);
        Bf16Problem q = pb;
        q.nMB = nfast;
        q.nNB = (int)(N / (256 * cb));
        q.mb0 = 0;
        q.sp0 = sp_first;
        q.nS = sp_count;
        const dim3 grid((unsigned)(q.nMB * q.nNB * q.nS));
        const bool gs = q.g.gauss != 0;
        if (cb == BF_CB) launch_pair<BF_CB>(n_terms, gs, grid, q, stream);
        else if (cb == BF_CB / 2) launch_pair<(BF_CB / 2 > 0 ? BF_CB / 2 : 1)>(n_terms, gs, grid, q, stream);
        else launch_pair<(BF_CB / 4 > 0 ? BF_CB / 4 : 1)>(n_terms, gs, grid, q, stream);
    };
    // K-range pipeline: when every row block runs on the two-wave kernel, the
    // split of later K ranges runs on the workspace's helper stream while the
    // main kernel multiplies earlier ones (the split's lean waves fit beside
    // the main kernel's on each CU); the caller's stream waits for each range's
    // split before that range's product, so all helper work is joined before
    // the reduction.
    const int64_t tiles_per_split = pb.kw / SPL_KT;
    const bool pipelined = ars == 1 && nfast > 0 && nMB_all == nfast64 && pb.nS >= 8 && pb.kw % SPL_KT == 0;
    // the pipeline also sums the first K ranges' partials on the helper
    // stream during the last range's product (FP64, M x N)
    const bool presum = pipelined && (M * N) % 4 == 0;
    const size_t plane_elems = (size_t)safe_mul(pb.Kp, pb.Np);
    const size_t planes_bytes = ((size_t)n_terms * plane_elems * 2 + 255) / 256 * 256;
    const size_t part_bytes = ((size_t)safe_mul(safe_mul(pb.nS, M), N) * 4 + 255) / 256 * 256;
    const size_t pre_bytes = presum ? (size_t)safe_mul(M, N) * 8 : 0;
    char *ws = (char *)scratch(SCRATCH_BF16, planes_bytes + part_bytes + pre_bytes, stream);
    pb.planes = (const uint16_t *)ws;
    pb.part = (float *)(ws + planes_bytes);
    double *pre = presum ? (double *)(ws + planes_bytes + part_bytes) : nullptr;
    int pre_splits = 0;  // splits already in pre
    if (pipelined) {
        // four equal K ranges: only the first range's split runs alone (at the
        // headline shape 7.34 ms per step against 7.62 without the pipeline;
        // an eighth then seven eighths measured 7.38 ms)
        constexpr int NCH = 4;
        int sb[NCH + 1];
        for (int c = 0; c <= NCH; ++c) sb[c] = (int)((int64_t)pb.nS * c / NCH);
        auto kt_of = [&](int sp) { return (int64_t)sp * tiles_per_split < kt_total ? (int64_t)sp * tiles_per_split
                                                                                   : kt_total; };
        hipStream_t hs = ws_aux_stream(stream);
        split_kunit(kt_of(sb[0]), kt_of(sb[1]) - kt_of(sb[0]), stream);
        hipEvent_t fork = ws_event(stream, 0);
        RBX_HIP(hipEventRecord(fork, stream));
        RBX_HIP(hipStreamWaitEvent(hs, fork, 0));
        for (int c = 1; c < NCH; ++c) {
            split_kunit(kt_of(sb[c]), kt_of(sb[c + 1]) - kt_of(sb[c]), hs);
            RBX_HIP(hipEventRecord(ws_event(stream, c), hs));
        }
        for (int c = 0; c < NCH; ++c) {
            if (c > 0) RBX_HIP(hipStreamWaitEvent(stream, ws_event(stream, c), 0));
            main_fast(sb[c], sb[c + 1] - sb[c]);
            if (presum && c == NCH - 2) {
                // partials of ranges 0 .. NCH - 2 are complete: sum them on
                // the helper stream while the last range is multiplied
                hipEvent_t ready = ws_event(stream, NCH);
                RBX_HIP(hipEventRecord(ready, stream));
                RBX_HIP(hipStreamWaitEvent(hs, ready, 0));
                pre_splits = sb[NCH - 1];
                ProfScope prof("bf16_presum_kernel", hs);
                hipLaunchKernelGGL(bf16_presum_kernel, dim3((unsigned)num_cus() * 4), dim3(256), 0, hs,
                                   (const float *)pb.part, pre_splits, M * N, pre);
                RBX_HIP(hipGetLastError());
                RBX_HIP(hipEventRecord(ws_event(stream, NCH + 1), hs));
            }
        }
        if (presum) RBX_HIP(hipStreamWaitEvent(stream, ws_event(stream, NCH + 1), 0));
    } else {
        if (ars == 1) {
            split_kunit(0, kt_total, stream);
        } else {
            int64_t total = (pb.Np / 32) * (pb.Kp / 16) * 64;  // one wave per (block, K step)
            int64_t blocks = (total + 255) / 256;
            if (blocks > 65536) blocks = 65536;
            ProfScope prof("bf16_split_kernel", stream);
            hipLaunchKernelGGL(split_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, A, ars, acs, K, N, pb.Kp,
                               pb.Np, n_terms, (uint16_t *)pb.planes);
            RBX_HIP(hipGetLastError());
        }
        if (nfast > 0) main_fast(0, pb.nS);
        if (nMB_all > nfast64) {
            ProfScope prof("bf16_skge_kernel", stream);
            Bf16Problem q = pb;
            q.nMB = nMB_all - nfast64;
            q.mb0 = nfast64;
            dim3 grid((unsigned)(q.nMB * q.nNB * q.nS));
            if (n_terms == 1) hipLaunchKernelGGL((bf16_skge_kernel<1, false>), grid, dim3(BF_NTHR), 0, stream, q);
            else if (n_terms == 2) hipLaunchKernelGGL((bf16_skge_kernel<2, false>), grid, dim3(BF_NTHR), 0, stream, q);
            else hipLaunchKernelGGL((bf16_skge_kernel<3, false>), grid, dim3(BF_NTHR), 0, stream, q);
            RBX_HIP(hipGetLastError());
        }
    }
    {
        const dim3 grid = reduce_grid(M, N);
        ProfScope prof("bf16_reduce_kernel", stream);
        hipLaunchKernelGGL((bf16_reduce_kernel<float, float>), grid, dim3(256), 0, stream, pb.part, pb.nS, M, N,
                           (double)alpha, (double)beta, B, brs, bcs, (const double *)pre, pre_splits);
        RBX_HIP(hipGetLastError());
    }
}

// Scratch of the wide-tile paths, or nullptr when the device cannot provide
// it: the caller then runs the tile kernels, which need none.
static void *wide_scratch(size_t bytes, hipStream_t stream) {
    try {
        return scratch(SCRATCH_BF16, bytes, stream);
    } catch (const Fail &f) {
        if (f.st != RBX_OUT_OF_MEMORY) throw;
        clear_error();
        return nullptr;
    }
}

// ---------------------------------------------------------------- wide FP32 tiles
// The plain FP32 fused sketch C = alpha * X * Y + beta * C with X the
// generated operator (M x K, rows along K) and Y the dense operand (K x N),
// for large interior shapes: the pair kernel's structure with FP32 operands.
// Y is packed once into FP32 fragment runs; one 512-thread workgroup per
// (64-row block, 256 CB-column block, K-split) generates each 32-deep
// operator chunk once into LDS (exact FP32 values, as the other kernels) and
// applies it to all of the block's columns with v_mfma_f32_32x32x2f32, so
// every operator entry is generated once per 1024 columns (CB = 4; 512 / 256
// for narrower outputs) instead of once per 256.  Partials go to scratch and
// are combined by bf16_reduce_kernel (FP64, split order).
//
// Packed Y: for 32-column block cb and 16-deep K step s, lane l's eight
// floats are Y[16 s + 2 q + l / 32][32 cb + l % 32], q = 0..7 (the B
// operands of the step's eight MFMAs), at ((cb * KS + s) * 64 + l) * 8 + q.
// The operator chunk in LDS uses the same order within each row: k = 16 st +
// kk sits at 16 st + 8 (kk % 2) + kk / 2.
constexpr int FW_LDK = 36;  // 32 k per row + 4 pad (floats)
// Largest packed operand + partials the wide kernels take in scratch; bigger
// problems run on the tile kernels, which need none.
constexpr size_t WIDE_SCRATCH_MAX = 16ull << 30;
static size_t align256b(size_t x) { return (x + 255) / 256 * 256; }

__device__ __forceinline__ int fw_pos(int k) { return (k & ~15) + 8 * (k & 1) + ((k & 15) >> 1); }

// Tail columns of the wide kernels: up to WIDE_TAIL output columns past the
// last whole tile (the right-hand side of [A b], a ragged edge) are formed by
// the threads that generate the operator, from the four entries each holds
// in registers per chunk, with VALU FMAs: the operator is generated once for
// the tiles and the tail together.  The generator of operator rows along K
// holds one row's four consecutive k (one accumulator per column); the one of
// counters along the rows holds four rows at one k (four per column).  The
// dense values of chunk c + 1 are loaded right after chunk c's are used, into
// the same registers, so they have a whole chunk's MFMAs to arrive (a copy
// from a second buffer would wait for the load at once).  TW
// columns are always loaded and accumulated (TW = 1, or 4 with the unused
// ones repeating the last real column), so the chunk loop has no branch: a
// load under a branch would make the compiler's memory-counter waits for
// the MFMA operands conservative.
constexpr int WIDE_TAIL = 4;
// The tail's columns: dense column j's element k at col[j][k * rs[j]], its
// output's element i at out[j][i * ors[j]]; entries nt .. WIDE_TAIL - 1 of
// col / rs repeat entry nt - 1.
template <typename T>
struct TailCols {
    const T *col[WIDE_TAIL];
    int64_t rs[WIDE_TAIL];
    T *out[WIDE_TAIL];
    int64_t ors[WIDE_TAIL];
    int nt;
};
template <typename T, bool ROWS, int TW>
struct WideTail {
    static constexpr int R = ROWS ? 4 : 1;  // output rows per thread
    static constexpr int E = ROWS ? 1 : 4;  // dense values per column per chunk
    T acc[TW][R];
    T cur[TW][E];  // the dense values of the chunk being generated

    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int j = 0; j < TW; ++j)
#pragma unroll
            for (int r = 0; r < R; ++r) acc[j][r] = (T)0;
    }
    // cur = dense values of the thread's k (k .. k + E - 1), every tail column
    __device__ __forceinline__ void load(const TailCols<T> &tc, int64_t k) {
#pragma unroll
        for (int j = 0; j < TW; ++j)
#pragma unroll
            for (int e = 0; e < E; ++e) cur[j][e] = tc.col[j][(k + e) * tc.rs[j]];
    }
    // acc += the chunk's operator entries (ent[0..3]) times cur
    __device__ __forceinline__ void step(const T (&ent)[4]) {
#pragma unroll
        for (int j = 0; j < TW; ++j) {
            if constexpr (ROWS) {
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[j][r] = fma(ent[r], cur[j][0], acc[j][r]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[j][0] = fma(ent[e], cur[j][e], acc[j][0]);
            }
        }
    }
    // The workgroup's 64 x nt partial into out[j * M + row], out = the
    // split's tail partials shifted to the block's first row.  Rows along K:
    // the eight threads of a row are adjacent lanes (butterfly).  Counters
    // along the rows: lanes l, l ^ 16, l ^ 32, l ^ 48 of a wave share rows
    // (butterfly), then the eight waves are summed in order through `red`
    // (8 * TW * 64 values of LDS free for reuse).  Every thread of the
    // workgroup calls this (it has a barrier).
    __device__ __forceinline__ void finish(int nt, int tid, int frow, int fcell, T *red, T *out, int64_t M) {
        const int lane = tid & 63, wave = tid >> 6;
        if constexpr (!ROWS) {
#pragma unroll
            for (int j = 0; j < TW; ++j) {
                T v = acc[j][0];
                v += __shfl_xor(v, 1);
                v += __shfl_xor(v, 2);
                v += __shfl_xor(v, 4);
                if (j < nt && fcell == 0) out[j * M + frow] = v;
            }
        } else {
#pragma unroll
            for (int j = 0; j < TW; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    T v = acc[j][r];
                    v += __shfl_xor(v, 16);
                    v += __shfl_xor(v, 32);
                    if (lane < 16) red[(wave * TW + j) * 64 + frow + r] = v;
                }
            __syncthreads();
            if (tid < 64 * nt) {
                const int j = tid >> 6, row = tid & 63;
                T v = red[j * 64 + row];
#pragma unroll
                for (int w = 1; w < BF_FTHR / 64; ++w) v += red[(w * TW + j) * 64 + row];
                out[j * M + row] = v;
            }
        }
    }
};

// The tail's outputs: out[j][i * ors[j]] = alpha * (sum over splits of
// tpart[(s * nt + j) * M + i]) + beta * out, FP64 sums.  G threads per output
// entry: G = 1 sums the splits in order; G = 64 (many splits) has lane l sum
// splits l, l + 64, ..., then a butterfly (a fixed order).
template <typename T, int G>
__global__ __launch_bounds__(256) void wide_tail_reduce_kernel(const T *tpart, int nS, int64_t M, TailCols<T> tc,
                                                               double alpha, double beta) {
    const int nt = tc.nt;
    const int g = threadIdx.x % G;
    const int64_t e = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;
    if (e >= M * nt) return;  // uniform per group
    const int64_t i = e % M;
    const int j = (int)(e / M);
    double v = 0.0;
    for (int s = g; s < nS; s += G) v += (double)tpart[((int64_t)s * nt + j) * M + i];
    if constexpr (G == 64) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (g != 0) return;
    }
    T *c = nullptr;
    int64_t rs = 0;
#pragma unroll
    for (int q = 0; q < WIDE_TAIL; ++q)
        if (q == j) {
            c = tc.out[q];
            rs = tc.ors[q];
        }
    c += i * rs;
    double out = alpha * v;
    if (beta != 0.0) out += beta * (double)(*c);
    *c = (T)out;
}

template <typename T>
static void launch_tail_reduce(const T *tpart, int nS, int64_t M, const TailCols<T> &tc, double alpha, double beta,
                               hipStream_t st) {
    const int64_t n = M * tc.nt;
    if (nS <= 16)
        hipLaunchKernelGGL((wide_tail_reduce_kernel<T, 1>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tpart,
                           nS, M, tc, alpha, beta);
    else
        hipLaunchKernelGGL((wide_tail_reduce_kernel<T, 64>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, tpart,
                           nS, M, tc, alpha, beta);
}


// How many of the columns past the whole tiles, [Nf, N), the wide kernel
// forms in its tail, the last nt of them (0: none; the others go to the
// caller as a rest problem), or -1 when the problem cannot run on the wide
// kernel.  The widest tiles (cb = 4) have no registers to spare for a tail.
// A problem whose last column lives apart (xD / xC) needs that column in the
// tail and whole 64-row blocks (a rest problem cannot carry it).
static int wide_tail_cols(const GemmProblem &pb, int64_t Mf, int64_t Nf, int cb) {
    const int64_t rem = pb.N - Nf;
    if (pb.xD != nullptr || pb.xC != nullptr) {
        if (pb.M != Mf || cb == 4 || rem < 1) return -1;
        return rem <= WIDE_TAIL ? (int)rem : 1;
    }
    return (rem >= 1 && rem <= WIDE_TAIL && cb < 4) ? (int)rem : 0;
}

// The tail's columns N - nt .. N - 1: dense column c at D[k * drs + c * dcs],
// output at C[i * crs + c * ccs], the last one at xD / xC when given.
template <typename T>
static TailCols<T> wide_tail_io(const GemmProblem &pb, int nt) {
    TailCols<T> tc{};
    tc.nt = nt;
    for (int j = 0; j < WIDE_TAIL; ++j) {
        if (j >= nt) {  // unused slots repeat the last column (loaded, never written)
            tc.col[j] = tc.col[nt - 1];
            tc.rs[j] = tc.rs[nt - 1];
            tc.out[j] = nullptr;
            tc.ors[j] = 0;
            continue;
        }
        const int64_t c = pb.N - nt + j;
        const bool last = c == pb.N - 1;
        tc.col[j] = (last && pb.xD) ? (const T *)pb.xD : (const T *)pb.D + c * pb.dcs;
        tc.rs[j] = (last && pb.xD) ? pb.xdrs : pb.drs;
        tc.out[j] = (last && pb.xC) ? (T *)pb.xC : (T *)pb.C + c * pb.ccs;
        tc.ors[j] = (last && pb.xC) ? pb.xcrs : pb.crs;
    }
    return tc;
}

// Unit-stride K: one workgroup per (32-column block, 256-deep K range), the
// tile staged through LDS (as split_kunit_kernel).
template <bool VEC>
__global__ __launch_bounds__(256) void pack_f32_kunit_kernel(const float *A, int64_t cs, int64_t K, int64_t N,
                                                             int64_t Kp, float *packed) {
    __shared__ __attribute__((aligned(16))) float tile[32][SPL_LD];
    const int tid = threadIdx.x;
    const int64_t KS = Kp / 16;
    const int64_t cb = blockIdx.y;
    const int64_t kt0 = (int64_t)blockIdx.x * SPL_KT;
    const int q4 = tid & 63, c0 = tid >> 6;
    const int64_t k = kt0 + 4 * q4;
    float4 v[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const int64_t j = cb * 32 + c0 + 4 * p;
        const float *src = A + k + j * cs;
        if (VEC && j < N && k + 3 < K) {
            v[p] = *(const float4 *)src;
        } else {
            v[p].x = (j < N && k < K) ? src[0] : 0.0f;
            v[p].y = (j < N && k + 1 < K) ? src[1] : 0.0f;
            v[p].z = (j < N && k + 2 < K) ? src[2] : 0.0f;
            v[p].w = (j < N && k + 3 < K) ? src[3] : 0.0f;
        }
    }
#pragma unroll
    for (int p = 0; p < 8; ++p) *(float4 *)&tile[c0 + 4 * p][4 * q4] = v[p];
    __syncthreads();
    const int64_t steps = (KS - kt0 / 16) < 16 ? (KS - kt0 / 16) : 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int task = tid + 256 * i;
        const int st = task >> 6, l = task & 63;
        if (st >= steps) break;
        const float *src = &tile[l & 31][16 * st + (l >> 5)];
        float4 lo, hi;
        lo.x = src[0]; lo.y = src[2]; lo.z = src[4]; lo.w = src[6];
        hi.x = src[8]; hi.y = src[10]; hi.z = src[12]; hi.w = src[14];
        float *dst = packed + ((cb * KS + kt0 / 16 + st) * 64 + l) * 8;
        *(float4 *)dst = lo;
        *(float4 *)(dst + 4) = hi;
    }
}

// Any strides: one thread per packed float.
__global__ __launch_bounds__(256) void pack_f32_kernel(const float *A, int64_t rs, int64_t cs, int64_t K, int64_t N,
                                                       int64_t Kp, int64_t Np, float *packed) {
    const int64_t KS = Kp / 16, total = (Np / 32) * KS * 512;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(e & 7), l = (int)((e >> 3) & 63);
        const int64_t run = e >> 9, s = run % KS, cb = run / KS;
        const int64_t k = 16 * s + 2 * q + (l >> 5), j = 32 * cb + (l & 31);
        packed[e] = (k < K && j < N) ? A[k * rs + j * cs] : 0.0f;
    }
}

struct F32WideProblem {
    int64_t M, N, Kp;  // M, N multiples of 64 / the tile width; Kp of the 32-deep chunk
    GenMap g;
    const float *packed;
    float *part;   // nS x M x N, row-major
    int64_t kw;    // K per split (multiple of 32)
    int nMB, nNB, nS;
    // tail columns (TAIL instantiations), partials at
    // tpart[(split * tc.nt + j) * M + row]
    TailCols<float> tc;
    float *tpart;
};

template <bool GAUSS, bool ROWS, int CB, int TW>
__global__ __launch_bounds__(BF_FTHR, 1) void f32_wide_kernel(F32WideProblem pb) {
    constexpr int RB = 2, MT = 64, NT = 256 * CB;  // CB 32-column blocks per wave
    __shared__ __attribute__((aligned(16))) float Ss[2][MT * FW_LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = pb.nMB * pb.nNB * pb.nS;
    const int L = blockIdx.x;
    const int t = (W % 8 == 0) ? (L % 8) * (W / 8) + L / 8 : L;  // XCD-aware: one K-split per XCD run
    const int mb = t % pb.nMB;
    const int rest = t / pb.nMB;
    const int nb = rest % pb.nNB;
    const int sp = rest / pb.nNB;
    const int64_t m0 = (int64_t)mb * MT, n0 = (int64_t)nb * NT;
    const int64_t kbeg = (int64_t)sp * pb.kw;
    int64_t kend = kbeg + pb.kw;
    if (kend > pb.Kp) kend = pb.Kp;
    const int nchunks = (int)((kend - kbeg) / 32);
    const GenMap g = pb.g;

    f32x16 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int64_t wcol = n0 + wave * (32 * CB);
    const int64_t KS = pb.Kp / 16;
    const float *bbase = pb.packed + ((wcol >> 5) * KS * 64 + lane) * 8;
    const int64_t bblk = KS * 512;  // floats between column blocks
    const int64_t s_first = kbeg / 16, s_last = kend / 16 - 1;

    // generator, operator rows along K (!ROWS): thread owns cell (tid & 7) of
    // operator row tid / 8; counters along the rows (ROWS): thread owns rows
    // 4 (tid & 15) .. + 3 of operator column tid / 16 of the chunk
    const int frow = ROWS ? 4 * (tid & 15) : tid >> 3, fcell = ROWS ? tid >> 4 : tid & 7;
    const uint64_t foff = ROWS ? (uint64_t)((g.a_off + kbeg + fcell) * g.P + ((g.b_off + m0) >> 2) + (tid & 15))
                               : (uint64_t)((g.a_off + m0 + frow) * g.P + ((g.b_off + kbeg) >> 2) + fcell);
    const uint64_t fstep = ROWS ? 32 * (uint64_t)g.P : 8;  // counters per chunk
    // tail columns: only the first column block's workgroups form them
    // (cb < 4: the output has one column block, nb == 0)
    WideTail<float, ROWS, TW == 0 ? 1 : TW> tl;
    const int64_t tk = kbeg + (ROWS ? fcell : 4 * fcell);  // the thread's first k of chunk 0
    if constexpr (TW > 0) {
        tl.zero();
        tl.load(pb.tc, tk);
    }
    auto gen_chunk = [&](int c, float *lds) {
        const float4 v = gen4(GAUSS, ctr_add(g.ctr, foff + fstep * (uint64_t)c), g.k0, g.k1);
        const float ent[4] = {entry_value<float>(v.x, GAUSS), entry_value<float>(v.y, GAUSS),
                              entry_value<float>(v.z, GAUSS), entry_value<float>(v.w, GAUSS)};
        if constexpr (ROWS) {
            const int pos = fw_pos(fcell);
#pragma unroll
            for (int e = 0; e < 4; ++e) lds[(frow + e) * FW_LDK + pos] = ent[e];
        } else {
            float *row = lds + frow * FW_LDK;
            const int k0 = 4 * fcell;  // even: k0, k0 + 2 share the even half; k0 + 1, k0 + 3 the odd half
            float2 ev, od;
            ev.x = ent[0];
            od.x = ent[1];
            ev.y = ent[2];
            od.y = ent[3];
            *(float2 *)(row + fw_pos(k0)) = ev;
            *(float2 *)(row + fw_pos(k0 + 1)) = od;
        }
        if constexpr (TW > 0) {
            tl.step(ent);
            tl.load(pb.tc, tk + 32 * (int64_t)(c + 1 < nchunks ? c + 1 : c));  // the next chunk's
        }
    };

    // B operands of one step: lo = q 0..3, hi = q 4..7 of block j
    float4 blo[CB], bhi[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        blo[j] = *(const float4 *)(bbase + j * bblk + s_first * 512);
        bhi[j] = *(const float4 *)(bbase + j * bblk + s_first * 512 + 4);
    }

    auto step = [&](const float *S, int ks, int64_t s_next) {
        __builtin_amdgcn_s_setprio(1);
        float4 alo[RB], ahi[RB];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const float *src = S + (32 * i + (lane & 31)) * FW_LDK + 16 * ks + 8 * (lane >> 5);
            alo[i] = *(const float4 *)src;
            ahi[i] = *(const float4 *)(src + 4);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int j = 0; j < CB; ++j)
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(f4_get(alo[i], q), f4_get(blo[j], q), acc[i][j],
                                                                     0, 0, 0);
#pragma unroll
        for (int j = 0; j < CB; ++j) blo[j] = *(const float4 *)(bbase + j * bblk + s_next * 512);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int j = 0; j < CB; ++j)
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(f4_get(ahi[i], q), f4_get(bhi[j], q), acc[i][j],
                                                                     0, 0, 0);
#pragma unroll
        for (int j = 0; j < CB; ++j) bhi[j] = *(const float4 *)(bbase + j * bblk + s_next * 512 + 4);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
    };

    const bool late = (wave & 4) != 0;
    gen_chunk(0, Ss[0]);
    __syncthreads();
    if constexpr (TW == 0) {
        for (int c = 0; c < nchunks; ++c) {
            const float *S = Ss[c & 1];
            float *Sn = Ss[(c + 1) & 1];
            const int64_t s0 = s_first + 2 * c;
            const bool more = c + 1 < nchunks;
            if (more && !late) gen_chunk(c + 1, Sn);
            step(S, 0, s0 + 1);
            step(S, 1, s0 + 2 > s_last ? s_last : s0 + 2);
            if (more && late) gen_chunk(c + 1, Sn);
            __syncthreads();
        }
    } else {
        // with a tail: one loop per wave role and the last chunk peeled, so
        // no branch surrounds the tail's loads inside a loop (the compiler's
        // memory-counter waits for the MFMA operands stay exact)
        const int last = nchunks - 1;
        if (!late) {
            for (int c = 0; c < last; ++c) {
                gen_chunk(c + 1, Ss[(c + 1) & 1]);
                step(Ss[c & 1], 0, s_first + 2 * c + 1);
                step(Ss[c & 1], 1, s_first + 2 * c + 2);
                __syncthreads();
            }
        } else {
            for (int c = 0; c < last; ++c) {
                step(Ss[c & 1], 0, s_first + 2 * c + 1);
                step(Ss[c & 1], 1, s_first + 2 * c + 2);
                gen_chunk(c + 1, Ss[(c + 1) & 1]);
                __syncthreads();
            }
        }
        step(Ss[last & 1], 0, s_first + 2 * last + 1);
        step(Ss[last & 1], 1, s_last);
        __syncthreads();
    }

    if constexpr (TW > 0)
        if (nb == 0) tl.finish(pb.tc.nt, tid, frow, fcell, &Ss[0][0], pb.tpart + (int64_t)sp * pb.tc.nt * pb.M + m0, pb.M);

    float *P = pb.part + (int64_t)sp * pb.M * pb.N;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = m0 + i * 32 + out_row32(lane, r);
                const int64_t gj = wcol + j * 32 + (lane & 31);
                P[gi * pb.N + gj] = acc[i][j][r];
            }
}

// Launches a wide-tile FP32 kernel instantiation by (family, orientation).
template <int CB, int TW>
static void launch_f32_wide(bool gauss, bool rows, dim3 grid, const F32WideProblem &w, hipStream_t st) {
    if (gauss && rows) hipLaunchKernelGGL((f32_wide_kernel<true, true, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
    else if (gauss) hipLaunchKernelGGL((f32_wide_kernel<true, false, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
    else if (rows) hipLaunchKernelGGL((f32_wide_kernel<false, true, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
    else hipLaunchKernelGGL((f32_wide_kernel<false, false, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
}
// by tail width: none, one column, two to four
template <int CB>
static void launch_f32_wide_tw(bool gauss, bool rows, int nt, dim3 grid, const F32WideProblem &w, hipStream_t st) {
    if (nt == 0) launch_f32_wide<CB, 0>(gauss, rows, grid, w, st);
    else if (nt == 1) launch_f32_wide<CB, 1>(gauss, rows, grid, w, st);
    else launch_f32_wide<CB, WIDE_TAIL>(gauss, rows, grid, w, st);
}

// Runs the part of an FP32 fused-sketch problem that the wide kernel covers
// (rows [0, Mf) x columns [0, Nf), Mf and Nf multiples of 64 and 1024) and
// returns in rest[] the remaining sub-problems, for the caller's tile kernels
// (-1: the wide kernel does not apply and nothing ran).
int f32_wide_run(const GemmProblem &pb, GemmProblem rest[2], hipStream_t stream) {
    if (!pb.gen_is_x || pb.sbuf != nullptr || (pb.g.b_off & 3) != 0) return -1;
    if (pb.K < 1024 || pb.K % 32 != 0 || pb.M < 64 || pb.N < 256) return -1;
    // the widest tile (1024, 512 or 256 columns) the output holds; the
    // columns past the last whole tile run on the tile kernels
    const int cb = pb.N >= 1024 ? 4 : (pb.N >= 512 ? 2 : 1);
    const int64_t NT = 256 * cb;
    const int64_t Mf = pb.M / 64 * 64, Nf = pb.N / NT * NT, K = pb.K;
    const int nt = wide_tail_cols(pb, Mf, Nf, cb);
    if (nt < 0) return -1;
    F32WideProblem w{};
    w.M = Mf;
    w.N = Nf;
    w.Kp = K;
    w.g = pb.g;
    w.nMB = (int)(Mf / 64);
    w.nNB = (int)(Nf / NT);
    int64_t kw = 4096;
    const int64_t tiles = (int64_t)w.nMB * w.nNB;
    const int64_t want = (4 * (int64_t)num_cus() + tiles - 1) / tiles;
    const int64_t kw_fill = (K + want - 1) / want;
    if (kw_fill < kw) kw = kw_fill;
    kw = (kw + 31) / 32 * 32;
    if (kw < 32) kw = 32;
    // short K over a large output: thin K-splits would cost more partial
    // traffic than the tile kernels' plain pass (measured at 4096 x 1024,
    // K = 1024)
    if (K < 2048 && kw < 256 && Mf * Nf > (1ll << 20)) return -1;
    w.kw = kw;
    w.nS = (int)((K + kw - 1) / kw);
    if ((int64_t)w.nMB * w.nNB * w.nS >= (1ll << 31)) return -1;
    const size_t packed_bytes = align256b((size_t)safe_mul(K, Nf) * 4);
    const size_t part_bytes = align256b((size_t)safe_mul(safe_mul(w.nS, Mf), Nf) * 4);
    const size_t tail_bytes = (size_t)w.nS * nt * Mf * 4;
    // the tile kernels need no scratch
    if (packed_bytes + part_bytes + tail_bytes > WIDE_SCRATCH_MAX) return -1;
    char *ws = (char *)wide_scratch(packed_bytes + part_bytes + tail_bytes, stream);
    if (ws == nullptr) return -1;
    w.packed = (const float *)ws;
    w.part = (float *)(ws + packed_bytes);
    const float *Y = (const float *)pb.D;
    w.tc = wide_tail_io<float>(pb, nt);
    w.tpart = (float *)(ws + packed_bytes + part_bytes);
    {
        ProfScope prof("f32_pack_kernel", stream);
        if (pb.drs == 1) {
            const int64_t kt = (K + SPL_KT - 1) / SPL_KT;
            RBX_REQUIRE(kt < (1ll << 31) && Nf / 32 <= 65535);
            const dim3 grid((unsigned)kt, (unsigned)(Nf / 32));
            if (pb.dcs % 4 == 0 && ((uintptr_t)Y & 15) == 0)
                hipLaunchKernelGGL(pack_f32_kunit_kernel<true>, grid, dim3(256), 0, stream, Y, pb.dcs, K, Nf, K,
                                   (float *)w.packed);
            else
                hipLaunchKernelGGL(pack_f32_kunit_kernel<false>, grid, dim3(256), 0, stream, Y, pb.dcs, K, Nf, K,
                                   (float *)w.packed);
        } else {
            hipLaunchKernelGGL(pack_f32_kernel, dim3(65536), dim3(256), 0, stream, Y, pb.drs, pb.dcs, K, Nf, K, Nf,
                               (float *)w.packed);
        }
        RBX_HIP(hipGetLastError());
    }
    {
        ProfScope prof("f32_wide_kernel", stream);
        const dim3 grid((unsigned)(w.nMB * w.nNB * w.nS));
        const bool rows = w.g.b_along_rows != 0;
        const bool gauss = w.g.gauss != 0;
        if (cb == 4) launch_f32_wide<4, 0>(gauss, rows, grid, w, stream);  // no tail at cb = 4
        else if (cb == 2) launch_f32_wide_tw<2>(gauss, rows, nt, grid, w, stream);
        else launch_f32_wide_tw<1>(gauss, rows, nt, grid, w, stream);
        RBX_HIP(hipGetLastError());
    }
    {
        ProfScope prof("f32_reduce_kernel", stream);
        const dim3 grid = reduce_grid(Mf, Nf);
        hipLaunchKernelGGL((bf16_reduce_kernel<float, float>), grid, dim3(256), 0, stream, (const float *)w.part, w.nS, Mf, Nf,
                           pb.alpha, pb.beta, (float *)pb.C, pb.crs, pb.ccs);
        if (nt > 0)
            launch_tail_reduce<float>(w.tpart, w.nS, Mf, w.tc, pb.alpha, pb.beta, stream);
        RBX_HIP(hipGetLastError());
    }
    int nr = 0;
    if (pb.M > Mf) {  // rows [Mf, M) x all columns: the operator's rows advance along a
        GemmProblem r = pb;
        r.M = pb.M - Mf;
        if (pb.g.b_along_rows) r.g.b_off = pb.g.b_off + Mf;
        else r.g.a_off = pb.g.a_off + Mf;
        r.C = (float *)pb.C + Mf * pb.crs;
        rest[nr++] = r;
    }
    if (pb.N - nt > Nf) {  // rows [0, Mf) x columns [Nf, N - nt)
        GemmProblem r = pb;
        r.xD = nullptr;
        r.xC = nullptr;
        r.M = Mf;
        r.N = pb.N - nt - Nf;
        r.D = Y + Nf * pb.dcs;
        r.C = (float *)pb.C + Nf * pb.ccs;
        rest[nr++] = r;
    }
    return nr;
}

// ---------------------------------------------------------------- wide FP64 tiles
// The plain FP64 fused sketch on the same plan as f32_wide_kernel, with
// v_mfma_f64_16x16x4f64: 64 x 128 CB tiles, 512 columns at CB = 4 (the FP64
// accumulators of a 64 x 1024 tile would not fit), so each operator entry is
// generated once per 512 columns (the tile kernels: once per 256).  Packed Y: for 16-column block jb
// and 16-deep step s, lane l's four doubles are Y[16 s + 4 q + l / 16][16 jb +
// l % 16], q = 0..3, at ((jb * KS + s) * 64 + l) * 4 + q; the operator chunk
// in LDS holds k = 16 st + kk at 16 st + 4 (kk % 4) + kk / 4.  Partials are
// FP64; K-splits are longer (FP64 chains need no bounding).
constexpr int DW_LDK = 36;  // doubles per LDS row: 32 k + 4 pad
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int dw_pos(int k) { return (k & ~15) + 4 * (k & 3) + ((k & 15) >> 2); }

__global__ __launch_bounds__(256) void pack_f64_kernel(const double *A, int64_t rs, int64_t cs, int64_t K, int64_t N,
                                                       int64_t Kp, int64_t Np, double *packed) {
    const int64_t KS = Kp / 16, total = (Np / 16) * KS * 256;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(e & 3), l = (int)((e >> 2) & 63);
        const int64_t run = e >> 8, s = run % KS, jb = run / KS;
        const int64_t k = 16 * s + 4 * q + (l >> 4), j = 16 * jb + (l & 15);
        packed[e] = (k < K && j < N) ? A[k * rs + j * cs] : 0.0;
    }
}

// Unit-stride K (column-major A): one workgroup per (16-column block, 256-deep
// K range); each column's 256 doubles are 16 threads' contiguous 16-byte
// loads, staged through LDS and written as the block's 16 runs.
constexpr int DPK_LD = 258;  // doubles per LDS row (256 + pad)
template <bool VEC>
__global__ __launch_bounds__(256) void pack_f64_kunit_kernel(const double *A, int64_t cs, int64_t K, int64_t N,
                                                             int64_t Kp, double *packed) {
    __shared__ __attribute__((aligned(16))) double tile[16][DPK_LD];
    const int tid = threadIdx.x;
    const int64_t KS = Kp / 16;
    const int64_t jb = blockIdx.y;
    const int64_t kt0 = (int64_t)blockIdx.x * SPL_KT;
    const int c = tid >> 4, p0 = tid & 15;
    const int64_t j = jb * 16 + c;
    f64x2 v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int64_t k = kt0 + 2 * (16 * i + p0);
        const double *src = A + k + j * cs;
        if (VEC && j < N && k + 1 < K) {
            v[i] = *(const f64x2 *)src;
        } else {
            v[i][0] = (j < N && k < K) ? src[0] : 0.0;
            v[i][1] = (j < N && k + 1 < K) ? src[1] : 0.0;
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) *(f64x2 *)&tile[c][2 * (16 * i + p0)] = v[i];
    __syncthreads();
    const int64_t steps = (KS - kt0 / 16) < 16 ? (KS - kt0 / 16) : 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int task = tid + 256 * i;
        const int st = task >> 6, l = task & 63;
        if (st >= steps) break;
        const double *src = &tile[l & 15][16 * st + (l >> 4)];
        f64x2 lo, hi;
        lo[0] = src[0];
        lo[1] = src[4];
        hi[0] = src[8];
        hi[1] = src[12];
        double *dst = packed + ((jb * KS + kt0 / 16 + st) * 64 + l) * 4;
        *(f64x2 *)dst = lo;
        *(f64x2 *)(dst + 2) = hi;
    }
}

struct F64WideProblem {
    int64_t M, N, Kp;
    GenMap g;
    const double *packed;
    double *part;  // nS x M x N, row-major
    int64_t kw;
    int nMB, nNB, nS;
    TailCols<double> tc;  // tail columns, as in F32WideProblem
    double *tpart;
};


template <bool GAUSS, bool ROWS, int CB, int TW>
__global__ __launch_bounds__(BF_FTHR, 1) void f64_wide_kernel(F64WideProblem pb) {
    constexpr int RB = 4, MT = 64, NT = 128 * CB;  // CB 16-column blocks per wave
    __shared__ __attribute__((aligned(16))) double Ss[2][MT * DW_LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = pb.nMB * pb.nNB * pb.nS;
    const int L = blockIdx.x;
    const int t = (W % 8 == 0) ? (L % 8) * (W / 8) + L / 8 : L;
    const int mb = t % pb.nMB;
    const int rest = t / pb.nMB;
    const int nb = rest % pb.nNB;
    const int sp = rest / pb.nNB;
    const int64_t m0 = (int64_t)mb * MT, n0 = (int64_t)nb * NT;
    const int64_t kbeg = (int64_t)sp * pb.kw;
    int64_t kend = kbeg + pb.kw;
    if (kend > pb.Kp) kend = pb.Kp;
    const int nchunks = (int)((kend - kbeg) / 32);
    const GenMap g = pb.g;

    f64x4 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    const int64_t wcol = n0 + wave * (16 * CB);
    const int64_t KS = pb.Kp / 16;
    const double *bbase = pb.packed + ((wcol >> 4) * KS * 64 + lane) * 4;
    const int64_t bblk = KS * 256;  // doubles between 16-column blocks
    const int64_t s_first = kbeg / 16, s_last = kend / 16 - 1;

    const int frow = ROWS ? 4 * (tid & 15) : tid >> 3, fcell = ROWS ? tid >> 4 : tid & 7;
    const uint64_t foff = ROWS ? (uint64_t)((g.a_off + kbeg + fcell) * g.P + ((g.b_off + m0) >> 2) + (tid & 15))
                               : (uint64_t)((g.a_off + m0 + frow) * g.P + ((g.b_off + kbeg) >> 2) + fcell);
    const uint64_t fstep = ROWS ? 32 * (uint64_t)g.P : 8;
    // tail columns: only the first column block's workgroups form them
    // (cb < 4: the output has one column block, nb == 0)
    WideTail<double, ROWS, TW == 0 ? 1 : TW> tl;
    const int64_t tk = kbeg + (ROWS ? fcell : 4 * fcell);  // the thread's first k of chunk 0
    if constexpr (TW > 0) {
        tl.zero();
        tl.load(pb.tc, tk);
    }
    auto gen_chunk = [&](int c, double *lds) {
        const float4 v = gen4(GAUSS, ctr_add(g.ctr, foff + fstep * (uint64_t)c), g.k0, g.k1);
        const double ent[4] = {entry_value<double>(v.x, GAUSS), entry_value<double>(v.y, GAUSS),
                               entry_value<double>(v.z, GAUSS), entry_value<double>(v.w, GAUSS)};
        if constexpr (ROWS) {
            const int pos = dw_pos(fcell);
#pragma unroll
            for (int e = 0; e < 4; ++e) lds[(frow + e) * DW_LDK + pos] = ent[e];
        } else {
            double *row = lds + frow * DW_LDK;
            const int k0 = 4 * fcell;
            row[dw_pos(k0)] = ent[0];
            row[dw_pos(k0 + 1)] = ent[1];
            row[dw_pos(k0 + 2)] = ent[2];
            row[dw_pos(k0 + 3)] = ent[3];
        }
        if constexpr (TW > 0) {
            tl.step(ent);
            tl.load(pb.tc, tk + 32 * (int64_t)(c + 1 < nchunks ? c + 1 : c));  // the next chunk's
        }
    };

    // B operands of one step: lo = q 0..1, hi = q 2..3 of block j
    f64x2 blo[CB], bhi[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        blo[j] = *(const f64x2 *)(bbase + j * bblk + s_first * 256);
        bhi[j] = *(const f64x2 *)(bbase + j * bblk + s_first * 256 + 2);
    }

    auto step = [&](const double *S, int ks, int64_t s_next) {
        __builtin_amdgcn_s_setprio(1);
        f64x2 alo[RB], ahi[RB];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const double *src = S + (16 * i + (lane & 15)) * DW_LDK + 16 * ks + 4 * (lane >> 4);
            alo[i] = *(const f64x2 *)src;
            ahi[i] = *(const f64x2 *)(src + 2);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int j = 0; j < CB; ++j)
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(alo[i][q], blo[j][q], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < CB; ++j) blo[j] = *(const f64x2 *)(bbase + j * bblk + s_next * 256);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int j = 0; j < CB; ++j)
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ahi[i][q], bhi[j][q], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < CB; ++j) bhi[j] = *(const f64x2 *)(bbase + j * bblk + s_next * 256 + 2);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
    };

    const bool late = (wave & 4) != 0;
    gen_chunk(0, Ss[0]);
    __syncthreads();
    if constexpr (TW == 0) {
        for (int c = 0; c < nchunks; ++c) {
            const double *S = Ss[c & 1];
            double *Sn = Ss[(c + 1) & 1];
            const int64_t s0 = s_first + 2 * c;
            const bool more = c + 1 < nchunks;
            if (more && !late) gen_chunk(c + 1, Sn);
            step(S, 0, s0 + 1);
            step(S, 1, s0 + 2 > s_last ? s_last : s0 + 2);
            if (more && late) gen_chunk(c + 1, Sn);
            __syncthreads();
        }
    } else {
        // with a tail: one loop per wave role and the last chunk peeled, so
        // no branch surrounds the tail's loads inside a loop (the compiler's
        // memory-counter waits for the MFMA operands stay exact)
        const int last = nchunks - 1;
        if (!late) {
            for (int c = 0; c < last; ++c) {
                gen_chunk(c + 1, Ss[(c + 1) & 1]);
                step(Ss[c & 1], 0, s_first + 2 * c + 1);
                step(Ss[c & 1], 1, s_first + 2 * c + 2);
                __syncthreads();
            }
        } else {
            for (int c = 0; c < last; ++c) {
                step(Ss[c & 1], 0, s_first + 2 * c + 1);
                step(Ss[c & 1], 1, s_first + 2 * c + 2);
                gen_chunk(c + 1, Ss[(c + 1) & 1]);
                __syncthreads();
            }
        }
        step(Ss[last & 1], 0, s_first + 2 * last + 1);
        step(Ss[last & 1], 1, s_last);
        __syncthreads();
    }

    if constexpr (TW > 0)
        if (nb == 0) tl.finish(pb.tc.nt, tid, frow, fcell, &Ss[0][0], pb.tpart + (int64_t)sp * pb.tc.nt * pb.M + m0, pb.M);

    // element r of a 16 x 16 FP64 block: row l / 16 + 4 r, column l % 16
    double *P = pb.part + (int64_t)sp * pb.M * pb.N;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gi = m0 + i * 16 + (lane >> 4) + 4 * r;
                const int64_t gj = wcol + j * 16 + (lane & 15);
                P[gi * pb.N + gj] = acc[i][j][r];
            }
}

template <int CB, int TW>
static void launch_f64_wide(bool gauss, bool rows, dim3 grid, const F64WideProblem &w, hipStream_t st) {
    if (gauss && rows) hipLaunchKernelGGL((f64_wide_kernel<true, true, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
    else if (gauss) hipLaunchKernelGGL((f64_wide_kernel<true, false, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
    else if (rows) hipLaunchKernelGGL((f64_wide_kernel<false, true, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
    else hipLaunchKernelGGL((f64_wide_kernel<false, false, CB, TW>), grid, dim3(BF_FTHR), 0, st, w);
}
// by tail width: none, one column, two to four
template <int CB>
static void launch_f64_wide_tw(bool gauss, bool rows, int nt, dim3 grid, const F64WideProblem &w, hipStream_t st) {
    if (nt == 0) launch_f64_wide<CB, 0>(gauss, rows, grid, w, st);
    else if (nt == 1) launch_f64_wide<CB, 1>(gauss, rows, grid, w, st);
    else launch_f64_wide<CB, WIDE_TAIL>(gauss, rows, grid, w, st);
}

int f64_wide_run(const GemmProblem &pb, GemmProblem rest[2], hipStream_t stream) {
    if (!pb.gen_is_x || pb.sbuf != nullptr || (pb.g.b_off & 3) != 0) return -1;
    if (pb.K < 1024 || pb.K % 32 != 0 || pb.M < 64 || pb.N < 128) return -1;
    const int cb = pb.N >= 512 ? 4 : (pb.N >= 256 ? 2 : 1);
    const int64_t NT = 128 * cb;
    const int64_t Mf = pb.M / 64 * 64, Nf = pb.N / NT * NT, K = pb.K;
    const int nt = wide_tail_cols(pb, Mf, Nf, cb);
    if (nt < 0) return -1;
    F64WideProblem w{};
    w.M = Mf;
    w.N = Nf;
    w.Kp = K;
    w.g = pb.g;
    w.nMB = (int)(Mf / 64);
    w.nNB = (int)(Nf / NT);
    int64_t kw = 16384;
    const int64_t tiles = (int64_t)w.nMB * w.nNB;
    // about four workgroups per CU; exactly one per CU when the tiles divide
    // the CUs (one round, a quarter of the partial sums: 4% faster at 16 tiles)
    const int64_t cus = num_cus();
    const int64_t want = cus % tiles == 0 ? cus / tiles : (4 * cus + tiles - 1) / tiles;
    const int64_t kw_fill = (K + want - 1) / want;
    if (kw_fill < kw) kw = kw_fill;
    kw = (kw + 31) / 32 * 32;
    if (kw < 32) kw = 32;
    if (K < 2048 && kw < 256 && Mf * Nf > (1ll << 20)) return -1;  // as in f32_wide_run
    w.kw = kw;
    w.nS = (int)((K + kw - 1) / kw);
    if ((int64_t)w.nMB * w.nNB * w.nS >= (1ll << 31)) return -1;
    const size_t packed_bytes = align256b((size_t)safe_mul(K, Nf) * 8);
    const size_t part_bytes = align256b((size_t)safe_mul(safe_mul(w.nS, Mf), Nf) * 8);
    const size_t tail_bytes = (size_t)w.nS * nt * Mf * 8;
    if (packed_bytes + part_bytes + tail_bytes > WIDE_SCRATCH_MAX) return -1;
    char *ws = (char *)wide_scratch(packed_bytes + part_bytes + tail_bytes, stream);
    if (ws == nullptr) return -1;
    w.packed = (const double *)ws;
    w.part = (double *)(ws + packed_bytes);
    const double *Y = (const double *)pb.D;
    w.tc = wide_tail_io<double>(pb, nt);
    w.tpart = (double *)(ws + packed_bytes + part_bytes);
    {
        ProfScope prof("f64_pack_kernel", stream);
        if (pb.drs == 1) {
            const int64_t kt = (K + SPL_KT - 1) / SPL_KT;
            RBX_REQUIRE(kt < (1ll << 31) && Nf / 16 <= 65535);
            const dim3 grid((unsigned)kt, (unsigned)(Nf / 16));
            if (pb.dcs % 2 == 0 && ((uintptr_t)Y & 15) == 0)
                hipLaunchKernelGGL(pack_f64_kunit_kernel<true>, grid, dim3(256), 0, stream, Y, pb.dcs, K, Nf, K,
                                   (double *)w.packed);
            else
                hipLaunchKernelGGL(pack_f64_kunit_kernel<false>, grid, dim3(256), 0, stream, Y, pb.dcs, K, Nf, K,
                                   (double *)w.packed);
        } else {
            hipLaunchKernelGGL(pack_f64_kernel, dim3(65536), dim3(256), 0, stream, Y, pb.drs, pb.dcs, K, Nf, K, Nf,
                               (double *)w.packed);
        }
        RBX_HIP(hipGetLastError());
    }
    {
        ProfScope prof("f64_wide_kernel", stream);
        const dim3 grid((unsigned)(w.nMB * w.nNB * w.nS));
        const bool rows = w.g.b_along_rows != 0;
        const bool gauss = w.g.gauss != 0;
        if (cb == 4) launch_f64_wide<4, 0>(gauss, rows, grid, w, stream);  // no tail at cb = 4
        else if (cb == 2) launch_f64_wide_tw<2>(gauss, rows, nt, grid, w, stream);
        else launch_f64_wide_tw<1>(gauss, rows, nt, grid, w, stream);
        RBX_HIP(hipGetLastError());
    }
    {
        ProfScope prof("f64_reduce_kernel", stream);
        const dim3 grid = reduce_grid(Mf, Nf);
        hipLaunchKernelGGL((bf16_reduce_kernel<double, double>), grid, dim3(256), 0, stream, (const double *)w.part,
                           w.nS, Mf, Nf, pb.alpha, pb.beta, (double *)pb.C, pb.crs, pb.ccs);
        if (nt > 0)
            launch_tail_reduce<double>(w.tpart, w.nS, Mf, w.tc, pb.alpha, pb.beta, stream);
        RBX_HIP(hipGetLastError());
    }
    int nr = 0;
    if (pb.M > Mf) {
        GemmProblem r = pb;
        r.M = pb.M - Mf;
        if (pb.g.b_along_rows) r.g.b_off = pb.g.b_off + Mf;
        else r.g.a_off = pb.g.a_off + Mf;
        r.C = (double *)pb.C + Mf * pb.crs;
        rest[nr++] = r;
    }
    if (pb.N - nt > Nf) {
        GemmProblem r = pb;
        r.xD = nullptr;
        r.xC = nullptr;
        r.M = Mf;
        r.N = pb.N - nt - Nf;
        r.D = Y + Nf * pb.dcs;
        r.C = (double *)pb.C + Nf * pb.ccs;
        rest[nr++] = r;
    }
    return nr;
}

// ---------------------------------------------------------------- fast generator bounds
// Measures, over every 32-bit word, the distance of the hardware factors of
// the fast generator from the portable factors that define the operator:
//   out[4 b + 0] = max |s' - s|,  out[4 b + 1] = max |c' - c|   (w as w0)
//   out[4 b + 2] = max |r' - r| / r over r > 0                    (w as w1)
//   out[4 b + 3] = number of words with (r' == 0) != (r == 0)
// per workgroup b; the host takes the maxima.  Word w = i + stride * j for
// thread i of the grid, so every word is visited exactly once.
constexpr int FGB_BLOCKS = 16384, FGB_THREADS = 256;
__global__ __launch_bounds__(FGB_THREADS) void fastgen_bounds_kernel(double *out) {
    __shared__ double red[4][FGB_THREADS];
    const uint32_t stride = (uint32_t)FGB_BLOCKS * FGB_THREADS;  // 2^22
    const uint32_t i = blockIdx.x * FGB_THREADS + threadIdx.x;
    double es = 0.0, ec = 0.0, er = 0.0;
    double zero_mismatch = 0.0;
    for (uint32_t j = 0; j < (1u << 10); ++j) {  // 2^22 * 2^10 = 2^32 words
        const uint32_t w = i + stride * j;
        float sf, cf;
        fast_sincos_factor(w, sf, cf);
        float s, c;
        rbm_sincos(__fmul_rn(3.1415926535897932f, uneg11(w)), s, c);
        es = fmax(es, fabs((double)sf - (double)s));
        ec = fmax(ec, fabs((double)cf - (double)c));
        const float rf = fast_radius_factor(w);
        const float r = __builtin_sqrtf(__fmul_rn(-2.0f, rbm_logf(u01(w))));
        if (r > 0.0f) er = fmax(er, fabs((double)rf - (double)r) / (double)r);
        if ((rf == 0.0f) != (r == 0.0f)) zero_mismatch += 1.0;
    }
    red[0][threadIdx.x] = es;
    red[1][threadIdx.x] = ec;
    red[2][threadIdx.x] = er;
    red[3][threadIdx.x] = zero_mismatch;
    __syncthreads();
    for (int h = FGB_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + h]);
            red[3][threadIdx.x] += red[3][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) out[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

static void fastgen_bounds(double measured[4], double assumed[2], hipStream_t stream) {
    RBX_REQUIRE(measured != nullptr && assumed != nullptr);
    double *dev_out = (double *)scratch(SCRATCH_SMALL, (size_t)FGB_BLOCKS * 4 * sizeof(double), stream);
    hipLaunchKernelGGL(fastgen_bounds_kernel, dim3(FGB_BLOCKS), dim3(FGB_THREADS), 0, stream, dev_out);
    RBX_HIP(hipGetLastError());
    std::vector<double> h((size_t)FGB_BLOCKS * 4);
    RBX_HIP(hipMemcpyAsync(h.data(), dev_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    RBX_HIP(hipStreamSynchronize(stream));
    for (int q = 0; q < 4; ++q) measured[q] = 0.0;
    for (int b = 0; b < FGB_BLOCKS; ++b) {
        for (int q = 0; q < 3; ++q) measured[q] = fmax(measured[q], h[4 * b + q]);
        measured[3] += h[4 * b + 3];
    }
    assumed[0] = (double)BM_ES;
    assumed[1] = (double)BM_ER;
}

}  // namespace rbx

using namespace rbx;

extern "C" rbx_status rbx_slskge3_bf16split(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                                            float alpha, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                                            const float *A, int64_t lda, float beta, float *B, int64_t ldb,
                                            int n_terms, void *stream) {
    RBX_API_BEGIN
    bf16_split_sketch(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb, n_terms,
                      (hipStream_t)stream);
    RBX_API_END
}

extern "C" rbx_status rbx_bf16_fastgen_bounds(double *measured, double *assumed, void *stream) {
    RBX_API_BEGIN
    fastgen_bounds(measured, assumed, (hipStream_t)stream);
    RBX_API_END
}
