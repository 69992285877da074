// rbx_skge.hip -- dense sketches lskge3 / rskge3 (skge.hh:154-355) and their
// adjoints, in FP64 and FP32, on the gfx950 matrix cores.
//
// Every call is reduced to one column-major product
//     C[M x N] = alpha * X[M x K] * Y[K x N] + beta * C
// in which exactly one of X, Y is the sketching operator G and the other is
// the dense data D, both addressed through (row stride, column stride).  A
// row-major call is the transposed column-major problem.  When the operator
// is not materialised (buff == NULL) each K-slice of G is generated from its
// Philox counters straight into LDS; G never reaches HBM.
//
// Tile: 64 x 256 outputs per 256-thread workgroup when the operator is X
// (256 x 64 when it is Y), 4 waves in 2 x 2, K-slice 16,
// v_mfma_{f64,f32}_16x16x4.  Each generated operator entry feeds 256 outputs.
// K-slices are double-buffered in LDS: while the MFMAs of slice s run, the
// data tile of slice s+1 is in flight from memory into registers, and the
// operator tile of slice s+1 is generated into the other buffer.
#include "rbx_internal.hh"
#include "rbx_gemm_common.hh"

namespace rbx {

template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output element r of lane l sits at row l/16 + 4r, column l%16
    static __device__ __forceinline__ int out_row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma16<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // output element r of lane l sits at row 4*(l/16) + r, column l%16
    static __device__ __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int BK = 16, NTHR = 256;

// Thread-to-element map of a PR x QC tile loaded from a strided matrix:
// consecutive threads walk the unit-stride direction.  Element e of this
// thread is (pp0 + e*dpp, qq0 + e*dqq).
template <int PR, int QC, int NT>
struct TileMap {
    static constexpr int NE = PR * QC / NT;
    static_assert(NE >= 1 && NT % PR == 0 && NT % QC == 0, "tile does not fit the thread map");
    int pp0, qq0, dpp, dqq;
    __device__ TileMap(bool p_unit, int tid) {
        if (p_unit) {
            pp0 = tid % PR; qq0 = tid / PR; dpp = 0; dqq = NT / PR;
        } else {
            qq0 = tid % QC; pp0 = tid / QC; dqq = 0; dpp = NT / QC;
        }
    }
};

// Aligned fast path of the operator tile: every Philox cell of the tile is
// whole and inside the operator.  Orientation 1 (major axis along q): cell cc
// of tile row pp holds entries (pp, 4cc .. 4cc+3).  Orientation 2 (major axis
// along p): cell cc of tile column qq holds entries (4cc .. 4cc+3, qq).  The
// thread's cells in the first K-slice are at counter offsets off0 + e*doff;
// each further K-slice adds `step`.  With fewer cells than threads, threads
// [0, NCELLS) generate one cell each and the others none.
template <int PR, int QC, int NT>
struct GenCells {
    static constexpr int NCELLS = PR * QC / 4;
    static_assert(NCELLS % NT == 0 || NT % NCELLS == 0, "operator tile does not fit the generator map");
    static constexpr int NG = NCELLS >= NT ? NCELLS / NT : 1;
    static constexpr int NTG = NCELLS >= NT ? NT : NCELLS;  // generating threads
    int a, cc;       // tile row (orientation 1) or column (orientation 2) of cell 0; cell index
    int da;          // row / column step between the thread's cells
    uint64_t off0, doff, step;
};

// FAST: every tile of the launch is interior (M, N multiples of the tile on
// the launched blocks, K a multiple of BK) and the generated operator's cells
// are aligned (b_off % 4 == 0), so no bounds logic is compiled in.  Blocks are
// offset by (bx0, by0) so that a FAST launch over the interior and generic
// launches over the ragged edge strips can split one product.
// GM >= 0 fixes the generated operator's family and orientation at compile
// time (bit 0: Gaussian, bit 1: major axis along the tile's p axis), so the
// FAST loop body is one basic block; GM < 0 reads them from the problem.
// NW waves (4 or 8) in a 2 x NW/2 grid over the tile.  With 8 waves two
// share each SIMD (the accumulators halve); the operator tile's cells go to
// the first waves only, so on every SIMD one wave's generator work overlaps
// the other wave's matrix work.
template <typename T, bool GEN_X, bool S_MEM, int BM, int BN, bool FAST, int GM, int NW>
__global__ __launch_bounds__(64 * NW, 1) void skge_kernel(GemmProblem pb_in, int bx0, int by0) {
    const GemmProblem pb = split_view<T>(pb_in, blockIdx.z);
    constexpr int NT = 64 * NW;
    constexpr int LDX = BM + 1, LDY = BN + 1;
    constexpr int WN = BN / (NW / 2);          // columns per wave
    constexpr int MI = BM / 32, NJ = WN / 16;  // 16x16 blocks per wave
    // operator tile OR x OC and data tile DR x DC (rows p, columns q)
    constexpr int OR = GEN_X ? BM : BK, OC = GEN_X ? BK : BN;
    constexpr int DR = GEN_X ? BK : BM, DC = GEN_X ? BN : BK;
    // LDS strides (p, q) of the operator and data tiles: X is stored [k][i]
    // (ldp 1, ldq LDX), Y is stored [k][j] (ldp LDY, ldq 1)
    constexpr int OLP = GEN_X ? 1 : LDY, OLQ = GEN_X ? LDX : 1;
    constexpr int DLP = GEN_X ? LDY : 1, DLQ = GEN_X ? 1 : LDX;
    __shared__ T Xs[2][BK * LDX];
    __shared__ T Ys[2][BK * LDY];
    using M16 = Mfma16<T>;
    typename M16::acc_t acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = typename M16::acc_t{0, 0, 0, 0};

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t m0 = (int64_t)(blockIdx.x + bx0) * BM;
    const int64_t n0 = (int64_t)(blockIdx.y + by0) * BN;
    const int64_t nk = (pb.K + BK - 1) / BK;
    const GenMap g = pb.g;
    const bool g_gauss = GM >= 0 ? (GM & 1) != 0 : g.gauss != 0;
    const bool g_rows = GM >= 0 ? (GM & 2) != 0 : g.b_along_rows != 0;

    // ---- data tile: D(p, q) at D[p*drs + q*dcs]; the K axis is p when the
    // data is Y (GEN_X) and q when it is X
    const T *D = (const T *)pb.D;
    const int64_t dp_lim = GEN_X ? pb.K : pb.M, dq_lim = GEN_X ? pb.N : pb.K;
    const int64_t dp0 = GEN_X ? 0 : m0, dq0 = GEN_X ? n0 : 0;
    const TileMap<DR, DC, NT> dmap(pb.drs == 1, tid);
    const int64_t d_estride = (int64_t)dmap.dpp * pb.drs + (int64_t)dmap.dqq * pb.dcs;
    const int64_t d_kstride = GEN_X ? BK * pb.drs : BK * pb.dcs;
    const bool d_fixed_in = FAST || (GEN_X ? (n0 + BN <= pb.N) : (m0 + BM <= pb.M));
    const T *dptr = D + (dp0 + dmap.pp0) * pb.drs + (dq0 + dmap.qq0) * pb.dcs;
    T dreg[TileMap<DR, DC, NT>::NE];

    // ---- materialised operator tile (S_MEM): G(p, q) at sbuf[soff + p*srs + q*scs]
    const T *Sb = (const T *)pb.sbuf + pb.soff;
    const int64_t op_lim = GEN_X ? pb.M : pb.K, oq_lim = GEN_X ? pb.K : pb.N;
    const int64_t op0 = GEN_X ? m0 : 0, oq0 = GEN_X ? 0 : n0;
    const bool o_fixed_in = FAST || (GEN_X ? (m0 + BM <= pb.M) : (n0 + BN <= pb.N));
    T oreg[S_MEM ? TileMap<OR, OC, NT>::NE : 1];

    // ---- generated operator tile, aligned fast path
    GenCells<OR, OC, NT> gc;
    const bool gen_fast = FAST || (!S_MEM && (g.b_off & 3) == 0 && o_fixed_in);
    if (!S_MEM) {
        const bool o1 = !g_rows;
        const int cells_q = OC / 4, cells_p = OR / 4;
        if (o1) {
            gc.cc = tid % cells_q; gc.a = tid / cells_q; gc.da = NT / cells_q;
            gc.off0 = (uint64_t)((g.a_off + op0 + gc.a) * g.P + ((g.b_off + oq0) >> 2) + gc.cc);
            gc.doff = (uint64_t)(gc.da * g.P);
            gc.step = GEN_X ? (uint64_t)(BK / 4) : (uint64_t)(BK * g.P);
        } else {
            gc.cc = tid % cells_p; gc.a = tid / cells_p; gc.da = NT / cells_p;
            gc.off0 = (uint64_t)((g.a_off + oq0 + gc.a) * g.P + ((g.b_off + op0) >> 2) + gc.cc);
            gc.doff = (uint64_t)(gc.da * g.P);
            gc.step = GEN_X ? (uint64_t)(BK * g.P) : (uint64_t)(BK / 4);
        }
    }

    auto load_data = [&](int64_t s) {
        const int64_t k0 = s * BK;
        if (FAST || (d_fixed_in && k0 + BK <= pb.K)) {
            const T *p = dptr + s * d_kstride;
#pragma unroll
            for (int e = 0; e < TileMap<DR, DC, NT>::NE; ++e) dreg[e] = p[e * d_estride];
        } else if constexpr (!FAST) {
#pragma unroll
            for (int e = 0; e < TileMap<DR, DC, NT>::NE; ++e) {
                const int64_t p = dp0 + (GEN_X ? k0 : 0) + dmap.pp0 + e * dmap.dpp;
                const int64_t q = dq0 + (GEN_X ? 0 : k0) + dmap.qq0 + e * dmap.dqq;
                dreg[e] = (p < dp_lim && q < dq_lim) ? D[p * pb.drs + q * pb.dcs] : (T)0;
            }
        }
    };
    auto store_data = [&](int buf) {
        T *lds = GEN_X ? Ys[buf] : Xs[buf];
#pragma unroll
        for (int e = 0; e < TileMap<DR, DC, NT>::NE; ++e)
            lds[(dmap.pp0 + e * dmap.dpp) * DLP + (dmap.qq0 + e * dmap.dqq) * DLQ] = dreg[e];
    };
    // materialised operator: same register staging with the operator's strides
    const TileMap<OR, OC, NT> omap(pb.srs == 1, tid);
    auto load_op = [&](int64_t s) {
        const int64_t k0 = s * BK;
        const bool in = FAST || (o_fixed_in && k0 + BK <= pb.K);
#pragma unroll
        for (int e = 0; e < (S_MEM ? TileMap<OR, OC, NT>::NE : 0); ++e) {
            const int64_t p = op0 + (GEN_X ? 0 : k0) + omap.pp0 + e * omap.dpp;
            const int64_t q = oq0 + (GEN_X ? k0 : 0) + omap.qq0 + e * omap.dqq;
            oreg[e] = (in || (p < op_lim && q < oq_lim)) ? Sb[p * pb.srs + q * pb.scs] : (T)0;
        }
    };
    auto store_op = [&](int buf) {
        T *lds = GEN_X ? Xs[buf] : Ys[buf];
#pragma unroll
        for (int e = 0; e < (S_MEM ? TileMap<OR, OC, NT>::NE : 0); ++e)
            lds[(omap.pp0 + e * omap.dpp) * OLP + (omap.qq0 + e * omap.dqq) * OLQ] = oreg[e];
    };
    auto gen_op = [&](int64_t s, int buf) {
        T *lds = GEN_X ? Xs[buf] : Ys[buf];
        const int64_t k0 = s * BK;
        if (FAST || (gen_fast && k0 + BK <= pb.K)) {
            const bool gauss = g_gauss;
            if (GenCells<OR, OC, NT>::NTG < NT && tid >= GenCells<OR, OC, NT>::NTG) return;
#pragma unroll
            for (int e = 0; e < GenCells<OR, OC, NT>::NG; ++e) {
                const float4 v = gen4(gauss, ctr_add(g.ctr, gc.off0 + e * gc.doff + s * gc.step), g.k0, g.k1);
                const int a = gc.a + e * gc.da;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int b = 4 * gc.cc + t;
                    const T val = entry_value<T>(f4_get(v, t), gauss);
                    if (!g_rows) lds[a * OLP + b * OLQ] = val;
                    else lds[b * OLP + a * OLQ] = val;
                }
            }
        } else if constexpr (!FAST) {
            if (GEN_X) gen_tile<T, OR, OC, NT>(g, m0, k0, pb.M, pb.K, lds, OLP, OLQ, tid);
            else gen_tile<T, OR, OC, NT>(g, k0, n0, pb.K, pb.N, lds, OLP, OLQ, tid);
        }
    };

    // prologue: K-slice 0 into buffer 0
    load_data(0);
    if (S_MEM) load_op(0);
    else gen_op(0, 0);
    store_data(0);
    if (S_MEM) store_op(0);
    __syncthreads();

    auto mma = [&](int cur) {
        const T *xs = Xs[cur];
        const T *ys = Ys[cur];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            T a[MI], b[NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i) a[i] = xs[kr * LDX + wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NJ; ++j) b[j] = ys[kr * LDY + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[i][j] = M16::mma(a[i], b[j], acc[i][j]);
        }
    };
    // One K-slice: the next slice's global loads are in flight during this
    // slice's MFMAs; the next operator tile is generated into the other
    // buffer, which was last read before the previous barrier.
    auto slice = [&](int64_t s) {
        const int cur = (int)(s & 1);
        load_data(s + 1);
        if (S_MEM) load_op(s + 1);
        mma(cur);
        if (!S_MEM) gen_op(s + 1, cur ^ 1);
        store_data(cur ^ 1);
        if (S_MEM) store_op(cur ^ 1);
        __syncthreads();
    };
    // the last slice is peeled so that the loop body has no branches
    for (int64_t s = 0; s + 1 < nk; ++s) slice(s);
    mma((int)((nk - 1) & 1));

    // epilogue: element r of the lane's accumulator is row M16::out_row(lane, r),
    // column lane%16 of each 16x16 block (the f64 and f32 forms differ)
    T *Cp = (T *)pb.C;
    const T alpha = (T)pb.alpha, beta = (T)pb.beta;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t gi = m0 + wm * (BM / 2) + i * 16 + M16::out_row(lane, r);
                int64_t gj = n0 + wn * WN + j * 16 + (lane & 15);
                if (FAST || (gi < pb.M && gj < pb.N)) {
                    T *c = Cp + gi * pb.crs + gj * pb.ccs;
                    T v = alpha * acc[i][j][r];
                    if (pb.beta != 0.0) v = v + beta * (*c);
                    *c = v;
                }
            }
}

template <typename T, bool GEN_X, bool S_MEM, int TM, int TN>
static void launch_tiles(const GemmProblem &pb, hipStream_t stream) {
    const int64_t gx = (pb.M + TM - 1) / TM, gy = (pb.N + TN - 1) / TN;
    RBX_REQUIRE(gx < (1ll << 31) && gy <= 65535);
    // interior blocks take the FAST kernel when K and the generator allow it
    // (with split K every range starts at a multiple of kw, itself a multiple
    // of BK, so the alignment of the first range holds for all of them)
    const bool fast_ok = pb.K % BK == 0 && (pb.nsplit <= 1 || pb.kw % BK == 0) &&
                         (S_MEM || (pb.g.b_off & 3) == 0);
    int64_t fx = fast_ok ? pb.M / TM : 0, fy = fast_ok ? pb.N / TN : 0;
    if (fx > 0 && fy > 0) {
        const dim3 grid((unsigned)fx, (unsigned)fy, (unsigned)pb.nsplit);
        const int gm = S_MEM ? -1 : (pb.g.gauss ? 1 : 0) | (pb.g.b_along_rows ? 2 : 0);
        constexpr int NF = 8, TF = 64 * NF;  // two waves per SIMD on interior tiles
        if (gm == 0) hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, TN, true, 0, NF>), grid, dim3(TF), 0, stream, pb, 0, 0);
        else if (gm == 1) hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, TN, true, 1, NF>), grid, dim3(TF), 0, stream, pb, 0, 0);
        else if (gm == 2) hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, TN, true, 2, NF>), grid, dim3(TF), 0, stream, pb, 0, 0);
        else if (gm == 3) hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, TN, true, 3, NF>), grid, dim3(TF), 0, stream, pb, 0, 0);
        else hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, TN, true, -1, NF>), grid, dim3(TF), 0, stream, pb, 0, 0);
    } else {
        fx = fy = 0;
    }
    // ragged edges (or everything) with the bounds-checked kernel: rows
    // [fx*TM, M) x all columns, then rows [0, fx*TM) x columns [fy*TN, N).
    // The long side of a tile (TN = 256 for a generated left operand, TM =
    // 256 for a generated right one) shrinks to 64 on an edge strip, so a
    // narrow remainder (n = 257, or a sketched vector) does not pay for a
    // whole 256-wide tile.
    constexpr int E = 64;
    static_assert(TM % E == 0 && TN % E == 0, "tiles are multiples of 64");
    const unsigned nz = (unsigned)pb.nsplit;
    if (gx > fx) {
        if (GEN_X && pb.N < TN) {
            hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, E, false, -1, 4>),
                               dim3((unsigned)(gx - fx), (unsigned)((pb.N + E - 1) / E), nz), dim3(NTHR), 0, stream,
                               pb, (int)fx, 0);
        } else if (!GEN_X && TM > E) {
            const int64_t r0 = fx * (TM / E);
            hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, E, TN, false, -1, 4>),
                               dim3((unsigned)((pb.M + E - 1) / E - r0), (unsigned)gy, nz), dim3(NTHR), 0, stream, pb,
                               (int)r0, 0);
        } else {
            hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, TN, false, -1, 4>),
                               dim3((unsigned)(gx - fx), (unsigned)gy, nz), dim3(NTHR), 0, stream, pb, (int)fx, 0);
        }
    }
    if (fx > 0 && gy > fy) {
        if (TN > E) {
            const int64_t c0 = fy * (TN / E);
            hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, E, false, -1, 4>),
                               dim3((unsigned)fx, (unsigned)((pb.N + E - 1) / E - c0), nz), dim3(NTHR), 0, stream, pb,
                               0, (int)c0);
        } else {
            hipLaunchKernelGGL((skge_kernel<T, GEN_X, S_MEM, TM, TN, false, -1, 4>),
                               dim3((unsigned)fx, (unsigned)(gy - fy), nz), dim3(NTHR), 0, stream, pb, 0, (int)fy);
        }
    }
    RBX_HIP(hipGetLastError());
}

// C = alpha * sum_z part[z] + beta * C, the sum in FP64 in split order.
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const T *part, int nsplit, int64_t M, int64_t N,
                                                            double alpha, double beta, T *C, int64_t crs,
                                                            int64_t ccs) {
    const int64_t total = M * N;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % M, j = idx / M;
        double v = 0.0;
        for (int z = 0; z < nsplit; ++z) v += (double)part[(int64_t)z * total + idx];
        T *c = C + i * crs + j * ccs;
        double out = alpha * v;
        if (beta != 0.0) out += beta * (double)(*c);
        *c = (T)out;
    }
}

template <typename T>
static void launch_skge_tiles(const GemmProblem &pb_in, hipStream_t stream) {
    const bool mem = pb_in.sbuf != nullptr;
    ProfScope prof(sizeof(T) == 8 ? "skge_kernel_f64" : "skge_kernel_f32", stream);
    GemmProblem pb = pb_in;
    // Split K when the output tiles alone do not cover the CUs: ranges of at
    // least 2048, enough of them for about two workgroups per CU.
    const int64_t TMo = pb.gen_is_x ? 64 : 256, TNo = pb.gen_is_x ? 256 : 64;
    const int64_t tiles = ((pb.M + TMo - 1) / TMo) * ((pb.N + TNo - 1) / TNo);
    const int64_t want = 2 * (int64_t)num_cus();
    if (tiles < num_cus() && pb.K >= 4096) {
        int64_t ns = (want + tiles - 1) / tiles;
        if (ns > pb.K / 2048) ns = pb.K / 2048;
        if (ns > 64) ns = 64;
        const int64_t kw = ((pb.K + ns - 1) / ns + BK - 1) / BK * BK;
        ns = (pb.K + kw - 1) / kw;
        if (ns > 1 && (uint64_t)ns * pb.M * pb.N * sizeof(T) <= (1ull << 30)) {
            pb.nsplit = (int)ns;
            pb.kw = kw;
            pb.part = scratch(SCRATCH_SPLITK, (size_t)ns * pb.M * pb.N * sizeof(T), stream);
        }
    }
    if (pb.gen_is_x) {
        if (!mem) launch_tiles<T, true, false, 64, 256>(pb, stream);
        else launch_tiles<T, true, true, 64, 256>(pb, stream);
    } else {
        if (!mem) launch_tiles<T, false, false, 256, 64>(pb, stream);
        else launch_tiles<T, false, true, 256, 64>(pb, stream);
    }
    if (pb.nsplit > 1) {
        int64_t blocks = (pb.M * pb.N + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, stream, (const T *)pb.part,
                           pb.nsplit, pb.M, pb.N, pb.alpha, pb.beta, (T *)pb.C, pb.crs, pb.ccs);
        RBX_HIP(hipGetLastError());
    }
}

// Large problems with a generated operator take the wide-tile kernels (each
// operator entry generated once per up to 1024 (FP32) / 512 (FP64) columns)
// for their interior and the tile kernels for the ragged rest.
// ---------------------------------------------------------------- one column
// C (M x 1) = alpha * X * y + beta * C with X the generated operator: a
// vector sketch.  The tile kernels would run their MFMAs over 255 empty
// columns; here each workgroup takes GV_ROWS rows and a K range of kw (a
// multiple of 4, chosen so the grid holds several workgroups per CU),
// every thread generates whole Philox cells of those rows (exact values, as
// everywhere) and accumulates v * y in FP64, a fixed-order workgroup
// reduction gives the range's partial, and gemv_reduce_kernel sums the ranges
// in order: deterministic, bound by the generator.  BR = 0: K runs along the
// counter's major index b (one cell = 4 consecutive k of one row); BR = 1: the
// rows run along b (one cell = 4 consecutive rows at one k).
constexpr int GV_ROWS = 8;

template <typename T, int BR>
__global__ __launch_bounds__(256) void gemv_gen_kernel(GenMap g, int64_t M, int64_t K, int64_t kw, const T *y,
                                                       int64_t ys, double *part) {
    __shared__ double red[GV_ROWS][256 / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * GV_ROWS, k0 = (int64_t)blockIdx.y * kw;
    int64_t k1 = k0 + kw;
    if (k1 > K) k1 = K;
    const bool gauss = g.gauss != 0;
    double acc[GV_ROWS];
#pragma unroll
    for (int r = 0; r < GV_ROWS; ++r) acc[r] = 0.0;
    if (BR == 0) {
        const int64_t c0 = k0 / 4, c1 = (k1 + 3) / 4;  // cells of this K range (k0 % 4 == 0)
        for (int64_t c = c0 + tid; c < c1; c += 256) {
            const int64_t k = 4 * c;
            T yv[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) yv[t] = k + t < K ? y[(k + t) * ys] : (T)0;
#pragma unroll
            for (int r = 0; r < GV_ROWS; ++r) {
                if (r0 + r >= M) break;
                const float4 v = gen4(gauss, ctr_add(g.ctr, (uint64_t)((g.a_off + r0 + r) * g.P + (g.b_off >> 2) + c)),
                                      g.k0, g.k1);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[r] += (double)entry_value<T>(f4_get(v, t), gauss) * (double)yv[t];
            }
        }
    } else {
        // b_off % 4 == 0 and r0 % 8 == 0: the block's rows are cells j = 0, 1
        const int64_t cb = (g.b_off + r0) >> 2;
        for (int64_t k = k0 + tid; k < k1; k += 256) {
            const double yv = (double)y[k * ys];
#pragma unroll
            for (int j = 0; j < GV_ROWS / 4; ++j) {
                if (r0 + 4 * j >= M) break;
                const float4 v = gen4(gauss, ctr_add(g.ctr, (uint64_t)((g.a_off + k) * g.P + cb + j)), g.k0, g.k1);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[4 * j + t] += (double)entry_value<T>(f4_get(v, t), gauss) * yv;
            }
        }
    }
    // fixed-order reduction: butterfly within each wave, then the waves in order
#pragma unroll
    for (int r = 0; r < GV_ROWS; ++r) {
        double v = acc[r];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[r][wave] = v;
    }
    __syncthreads();
    if (tid < GV_ROWS && r0 + tid < M) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) v += red[tid][w];
        part[(int64_t)blockIdx.y * M + r0 + tid] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void gemv_reduce_kernel(const double *part, int nk, int64_t M, double alpha,
                                                          double beta, T *C, int64_t crs) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int z = 0; z < nk; ++z) v += part[(int64_t)z * M + i];
        T *c = C + i * crs;
        double out = alpha * v;
        if (beta != 0.0) out += beta * (double)(*c);
        *c = (T)out;
    }
}

// Runs a one-column problem on gemv_gen_kernel; false when it does not apply.
template <typename T>
static bool gemv_run(const GemmProblem &pb, hipStream_t stream) {
    if (pb.N != 1 || !pb.gen_is_x || pb.sbuf != nullptr || (pb.g.b_off & 3) != 0) return false;
    if (pb.K < 1024) return false;
    const int64_t nb = (pb.M + GV_ROWS - 1) / GV_ROWS;
    if (nb >= (1ll << 31)) return false;
    // K ranges: enough for ~16 workgroups per CU, each 1024 to 16384 deep
    const int64_t want = (16 * (int64_t)num_cus() + nb - 1) / nb;
    int64_t kw = (pb.K + want - 1) / want;
    kw = (kw + 1023) / 1024 * 1024;
    if (kw > 16384) kw = 16384;
    const int64_t nk = (pb.K + kw - 1) / kw;
    if (nk > 65535) return false;
    double *part = (double *)scratch(SCRATCH_SPLITK, (size_t)safe_mul(nk, pb.M) * sizeof(double), stream);
    ProfScope prof("gemv_gen_kernel", stream);
    if (pb.g.b_along_rows)
        hipLaunchKernelGGL((gemv_gen_kernel<T, 1>), dim3((unsigned)nb, (unsigned)nk), dim3(256), 0, stream, pb.g, pb.M,
                           pb.K, kw, (const T *)pb.D, pb.drs, part);
    else
        hipLaunchKernelGGL((gemv_gen_kernel<T, 0>), dim3((unsigned)nb, (unsigned)nk), dim3(256), 0, stream, pb.g, pb.M,
                           pb.K, kw, (const T *)pb.D, pb.drs, part);
    int64_t blocks = (pb.M + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(gemv_reduce_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, stream, (const double *)part,
                       (int)nk, pb.M, pb.alpha, pb.beta, (T *)pb.C, pb.crs);
    RBX_HIP(hipGetLastError());
    return true;
}

// The same product described transposed: C^T = Y^T X^T, so a generated
// operator on the Y side becomes the X side (its orientation flag flips).
static GemmProblem transposed_problem(const GemmProblem &pb) {
    GemmProblem t = pb;
    t.M = pb.N;
    t.N = pb.M;
    t.gen_is_x = pb.gen_is_x ? 0 : 1;
    t.g.b_along_rows = pb.g.b_along_rows ? 0 : 1;
    t.drs = pb.dcs;
    t.dcs = pb.drs;
    t.crs = pb.ccs;
    t.ccs = pb.crs;
    if (pb.sbuf) {
        t.srs = pb.scs;
        t.scs = pb.srs;
    }
    return t;
}

template <typename T>
static void launch_skge(const GemmProblem &pb, hipStream_t stream) {
    GemmProblem rest[2];
    auto wide = [&](const GemmProblem &q) {
        const int nr = sizeof(T) == 4 ? f32_wide_run(q, rest, stream) : f64_wide_run(q, rest, stream);
        for (int i = 0; i < nr; ++i)
            if (!gemv_run<T>(rest[i], stream)) launch_skge_tiles<T>(rest[i], stream);
        return nr >= 0;
    };
    if (wide(pb)) return;
    if (!pb.gen_is_x && pb.sbuf == nullptr && wide(transposed_problem(pb))) return;
    if (gemv_run<T>(pb, stream)) return;
    if (!pb.gen_is_x && pb.sbuf == nullptr && pb.M == 1 && gemv_run<T>(transposed_problem(pb), stream)) return;
    launch_skge_tiles<T>(pb, stream);
}

// C = beta * C on an M x N strided matrix (util::safe_scal semantics).
template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(T *C, int64_t M, int64_t N, int64_t crs, int64_t ccs, double beta) {
    int64_t total = M * N;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = (crs == 1) ? idx % M : idx / N;
        int64_t j = (crs == 1) ? idx / M : idx % N;
        T *c = C + i * crs + j * ccs;
        *c = (beta == 0.0) ? (T)0 : (T)beta * (*c);
    }
}

template <typename T>
void scale_output(T *C, int64_t M, int64_t N, int64_t crs, int64_t ccs, double beta, hipStream_t stream) {
    if (M <= 0 || N <= 0 || beta == 1.0) return;
    int64_t blocks = (M * N + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    ProfScope prof("scale_kernel", stream);
    hipLaunchKernelGGL(scale_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, stream, C, M, N, crs, ccs, beta);
    RBX_HIP(hipGetLastError());
}
template void scale_output<float>(float *, int64_t, int64_t, int64_t, int64_t, double, hipStream_t);
template void scale_output<double>(double *, int64_t, int64_t, int64_t, int64_t, double, hipStream_t);

// Validate a left sketch call exactly as lskge3 does (skge.hh:174-193) and
// build the column-major problem.
template <typename T>
GemmProblem plan_left(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                      const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda,
                      double beta, T *B, int64_t ldb) {
    RBX_REQUIRE(S != nullptr);
    check_layout(layout);
    check_op(opS);
    check_op(opA);
    RBX_REQUIRE(d >= 0 && n >= 0 && m >= 0);
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
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
    GemmProblem pb{};
    pb.alpha = alpha;
    pb.beta = beta;
    pb.D = A;
    pb.C = B;
    // op(A)[k,j] strides in the column-major picture
    if (layout == 'C') {
        pb.M = d; pb.N = n; pb.K = m;
        pb.gen_is_x = 1;
        pb.g = gen_map(*S, opS, ro_s, co_s, false);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    } else {
        pb.M = n; pb.N = d; pb.K = m;
        pb.gen_is_x = 0;
        pb.g = gen_map(*S, opS, ro_s, co_s, true);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    }
    set_mem_operand(pb, *S, opS, ro_s, co_s, layout == 'R');
    return pb;
}

// rskge3 (skge.hh:326-346)
template <typename T>
GemmProblem plan_right(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                       const T *A, int64_t lda, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       double beta, T *B, int64_t ldb) {
    RBX_REQUIRE(S != nullptr);
    check_layout(layout);
    check_op(opS);
    check_op(opA);
    RBX_REQUIRE(d >= 0 && n >= 0 && m >= 0);
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
    DenseGeom geo = dense_geom(S->dist);
    safe_mul(geo.dim_minor, geo.P);
    int64_t rows_S = opS == 'N' ? n : d, cols_S = opS == 'N' ? d : n;
    RBX_REQUIRE(S->dist.n_rows >= rows_S + ro_s);
    RBX_REQUIRE(S->dist.n_cols >= cols_S + co_s);
    int64_t rows_A = opA == 'N' ? m : n, cols_A = opA == 'N' ? n : m;
    if (layout == 'C') {
        RBX_REQUIRE(lda >= rows_A);
        RBX_REQUIRE(ldb >= m);
    } else {
        RBX_REQUIRE(lda >= cols_A);
        RBX_REQUIRE(ldb >= d);
    }
    GemmProblem pb{};
    pb.alpha = alpha;
    pb.beta = beta;
    pb.D = A;
    pb.C = B;
    if (layout == 'C') {
        pb.M = m; pb.N = d; pb.K = n;
        pb.gen_is_x = 0;
        pb.g = gen_map(*S, opS, ro_s, co_s, false);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    } else {
        pb.M = d; pb.N = m; pb.K = n;
        pb.gen_is_x = 1;
        pb.g = gen_map(*S, opS, ro_s, co_s, true);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    }
    set_mem_operand(pb, *S, opS, ro_s, co_s, layout == 'R');
    return pb;
}

template GemmProblem plan_left<float>(char, char, char, int64_t, int64_t, int64_t, double, const rbx_dense_skop *,
                                      int64_t, int64_t, const float *, int64_t, double, float *, int64_t);

template <typename T>
static void run_problem(const GemmProblem &pb, hipStream_t stream) {
    if (pb.M == 0 || pb.N == 0) return;
    if (pb.alpha == 0.0 || pb.K == 0) {
        scale_output<T>((T *)pb.C, pb.M, pb.N, pb.crs, pb.ccs, pb.beta, stream);
        return;
    }
    launch_skge<T>(pb, stream);
}

template <typename T>
bool left_sketch_with_column(char opS, int64_t d, int64_t n, int64_t m, const rbx_dense_skop *S, int64_t ro_s,
                             int64_t co_s, const T *A, int64_t lda, const T *a, T *B, int64_t ldb, T *b,
                             hipStream_t stream) {
    if (d == 0 || m == 0) return false;
    GemmProblem pb = plan_left<T>('C', opS, 'N', d, n + 1, m, 1.0, S, ro_s, co_s, A, lda, 0.0, B, ldb);
    pb.xD = a;
    pb.xdrs = 1;
    pb.xC = b;
    pb.xcrs = 1;
    GemmProblem rest[2];
    const int nr = sizeof(T) == 4 ? f32_wide_run(pb, rest, stream) : f64_wide_run(pb, rest, stream);
    if (nr < 0) return false;
    for (int i = 0; i < nr; ++i) launch_skge<T>(rest[i], stream);  // columns between the tiles and the tail
    return true;
}
template bool left_sketch_with_column<float>(char, int64_t, int64_t, int64_t, const rbx_dense_skop *, int64_t,
                                             int64_t, const float *, int64_t, const float *, float *, int64_t,
                                             float *, hipStream_t);
template bool left_sketch_with_column<double>(char, int64_t, int64_t, int64_t, const rbx_dense_skop *, int64_t,
                                              int64_t, const double *, int64_t, const double *, double *, int64_t,
                                              double *, hipStream_t);

}  // namespace rbx

using namespace rbx;

extern "C" {

rbx_status rbx_slskge3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const float *A, int64_t lda,
                       float beta, float *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<float>(plan_left<float>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dlskge3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const double *A, int64_t lda,
                       double beta, double *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<double>(plan_left<double>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb),
                        (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_srskge3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, float alpha,
                       const float *A, int64_t lda, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       float beta, float *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<float>(plan_right<float>(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_drskge3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                       const double *A, int64_t lda, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       double beta, double *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<double>(plan_right<double>(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb),
                        (hipStream_t)stream);
    RBX_API_END
}

// Adjoint of B = op(S) * A with respect to A: Abar = op(S)^T * Bbar, i.e. the
// same generated operator read with the opposite op.
rbx_status rbx_slskge3_adjoint(char layout, char opS, int64_t d, int64_t n, int64_t m, float alpha,
                               const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const float *Bbar,
                               int64_t ldbb, float beta, float *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<float>(plan_left<float>(layout, opT, 'N', m, n, d, alpha, S, ro_s, co_s, Bbar, ldbb, beta, Abar, ldab),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dlskge3_adjoint(char layout, char opS, int64_t d, int64_t n, int64_t m, double alpha,
                               const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const double *Bbar,
                               int64_t ldbb, double beta, double *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<double>(plan_left<double>(layout, opT, 'N', m, n, d, alpha, S, ro_s, co_s, Bbar, ldbb, beta, Abar, ldab),
                        (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_srskge3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, float alpha,
                               const float *Bbar, int64_t ldbb, const rbx_dense_skop *S, int64_t ro_s,
                               int64_t co_s, float beta, float *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<float>(plan_right<float>(layout, 'N', opT, m, n, d, alpha, Bbar, ldbb, S, ro_s, co_s, beta, Abar, ldab),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_drskge3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                               const double *Bbar, int64_t ldbb, const rbx_dense_skop *S, int64_t ro_s,
                               int64_t co_s, double beta, double *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<double>(plan_right<double>(layout, 'N', opT, m, n, d, alpha, Bbar, ldbb, S, ro_s, co_s, beta, Abar, ldab),
                        (hipStream_t)stream);
    RBX_API_END
}

}  // extern "C"
