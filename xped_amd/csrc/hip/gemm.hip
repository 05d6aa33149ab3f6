// Grouped, persistent block GEMM on the CDNA4 matrix cores.
//
// The reference multiplies matched sectors one at a time with Eigen/BLAS
// (src/Core/TensorBase.cpp:262-279 -> MatrixInterface::prod, MatrixInterface_Eigen_impl.cpp:155-160).
// Here every (sector, output tile, K chunk) triple of one contraction is an entry of a work list
// processed by one persistent launch (FP64: sectors with both output sides >= 192 in 128 x 128
// tiles by one launch, the others in 64 x 64 tiles by a second).  The list is cut into 8 contiguous parts
// of equal work (sum of K over the tiles), one per XCD: workgroup b serves part b % 8, so the
// tiles of one sector run on one XCD and share its L2, and uneven sector sizes are balanced.
// Launches with few output tiles (a sketch's Gram matrix: a small output with a long K) split K
// into chunks that write FP64 partial products to a workspace; a second kernel sums the chunks
// in a fixed order (deterministic) and applies alpha / beta.  Items with one short output side
// (<= 32) and a short K (<= 32), as the contractions of a site tensor's physical or virtual leg
// with a large environment tensor give, are memory-bound and run on a third kernel (one thread per
// index of the long side, FP64 FMA) in the same call.
//
// Precision modes (operands are always FP64 in memory, the output is FP64):
//   XH_FP64   v_mfma_f64_16x16x4_f64, FP64 accumulation
//   XH_FP32   operands rounded to FP32, v_mfma_f32_16x16x4_f32, FP32 accumulation
//   XH_BF16X3 operands rounded to FP32 and split x = hi + lo with hi, lo in BF16;
//             A*B ~ hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x16_bf16, FP32 accumulation
//   XH_OZAKI  FP64 accuracy from the int8 matrix cores (error-free splitting, Ozaki scheme):
//             every row of A and column of B is scaled by a power of two to |x| < 1 and cut into
//             OZ_NS slices of 7 bits (int8); the slice products s + t <= OZ_NS + 1 run on
//             v_mfma_i32_16x16x64_i8 with exact int32 sums, and are combined in FP64
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_map>

#include "device.hpp"
#include "../common/pool.hpp"

using xh::check;
using xh::guard;

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64;
constexpr int G_THREADS = 256;

constexpr int NPART = 8;  // XCDs of an MI355X

struct Tile {
    int32_t item, m0, n0, ks;  // ks: K chunk
};

// K split of an item: chunks of kc (the last one shorter); split > 1 writes partials to ws.
struct ItemExt {
    int32_t split, kc;
    int64_t ws_off;
};

struct GemmArgs {
    const xh_gemm_item* items;
    const ItemExt* ext;
    const Tile* tiles;
    int32_t part[NPART + 1];  // tile range of each XCD's part
    int32_t ta, tb;
    const double* A;
    const double* B;
    double* C;
    double* ws;
    double alpha, beta;
};

__device__ __forceinline__ double load_a(const GemmArgs& g, const xh_gemm_item& it, int i, int kk, int kend)
{
    if(i >= it.m || kk >= kend) return 0.0;
    return g.ta ? g.A[it.a_off + int64_t(i) * it.k + kk] : g.A[it.a_off + int64_t(kk) * it.m + i];
}
__device__ __forceinline__ double load_b(const GemmArgs& g, const xh_gemm_item& it, int kk, int j, int kend)
{
    if(kk >= kend || j >= it.n) return 0.0;
    return g.tb ? g.B[it.b_off + int64_t(kk) * it.n + j] : g.B[it.b_off + int64_t(j) * it.k + kk];
}

// The tile loop of workgroup blockIdx.x: its XCD's part, strided by the workgroups of that XCD.
struct TileRange {
    int32_t begin, end, stride;
};
__device__ __forceinline__ TileRange my_tiles(const GemmArgs& g)
{
    const int x = blockIdx.x % NPART;
    const int per = (gridDim.x + NPART - 1 - x) / NPART;  // workgroups on this XCD
    return TileRange{g.part[x] + int32_t(blockIdx.x / NPART), g.part[x + 1], per};
}

// Output layout of the 16x16 accumulators: for the FP32-accumulating MFMAs register r of lane l
// holds row 4*(l/16) + r; for v_mfma_f64_16x16x4_f64 it holds row (l/16) + 4*r.  Column = l % 16.
template <bool F64LAYOUT, typename Acc>
__device__ __forceinline__ void store_tile(const GemmArgs& g, const xh_gemm_item& it, const ItemExt& ex, const Tile& tl, int wm, int wn, int lane,
                                           Acc acc[2][2])
{
    // a K chunk of a split item stores its raw partial product into the workspace
    const bool part = ex.split > 1;
    double* base = part ? g.ws + ex.ws_off + int64_t(tl.ks) * it.m * it.n : g.C + it.c_off;
    const double alpha = part ? 1.0 : g.alpha, beta = part ? 0.0 : g.beta;
    for(int mi = 0; mi < 2; ++mi)
        for(int ni = 0; ni < 2; ++ni)
            for(int r = 0; r < 4; ++r) {
                const int rr = F64LAYOUT ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                const int row = tl.m0 + wm * 32 + mi * 16 + rr;
                const int col = tl.n0 + wn * 32 + ni * 16 + (lane & 15);
                if(row < it.m && col < it.n) {
                    double* p = base + int64_t(col) * it.m + row;
                    const double v = static_cast<double>(acc[mi][ni][r]);
                    *p = (beta == 0.0) ? alpha * v : alpha * v + beta * *p;
                }
            }
}

// C = alpha * sum of the K chunks + beta * C for the split items (blockIdx.y: split item).
struct RedItem {
    int64_t c_off, ws_off, mn;
    int32_t split, pad;
};
__global__ __launch_bounds__(256) void k_splitk_reduce(const RedItem* __restrict__ red, const double* __restrict__ ws, double* __restrict__ C, double alpha,
                                                       double beta)
{
    const RedItem r = red[blockIdx.y];
    for(int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < r.mn; e += int64_t(gridDim.x) * blockDim.x) {
        double acc = 0.0;
        for(int s = 0; s < r.split; ++s) acc += ws[r.ws_off + s * r.mn + e];
        double* p = C + r.c_off + e;
        *p = (beta == 0.0) ? alpha * acc : alpha * acc + beta * *p;
    }
}

// ---------------------------------------------------------------- FP64, pipelined
// BM x BN output tiles, 4 waves in 2 x 2, each owning (BM/2) x (BN/2) = MI x NI blocks of 16 x 16.
// K steps of 16 run through two LDS stages: the operands of step t+1 are fetched into registers
// while the matrix cores work on step t, then written to the other stage; one barrier per step.
// Tiles that lie inside their item (all rows, columns and the whole K step in range) load without
// bounds checks from per-thread offsets computed once per tile.
template <int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(64 * WGM * WGN, 4) void k_gemm_f64p(GemmArgs g)
{
    constexpr int NT = 64 * WGM * WGN;                 // threads
    constexpr int WTM = BM / WGM, WTN = BN / WGN;       // wave tile
    constexpr int MI = WTM / 16, NI = WTN / 16;
    constexpr int RA = BM * BK / NT, RB = BN * BK / NT;  // operand elements per thread and step
    __shared__ double sA[2][BK][BM + 1];
    __shared__ double sB[2][BK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const TileRange tr = my_tiles(g);
    for(int t = tr.begin; t < tr.end; t += tr.stride) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        const ItemExt ex = g.ext[tl.item];
        const int kbeg = tl.ks * ex.kc, kend = min(it.k, kbeg + ex.kc);
        const bool full_mn = tl.m0 + BM <= it.m && tl.n0 + BN <= it.n;
        // element r of this thread: A (i, kk) and B (kk, j), lanes along each operand's contiguous direction
        int ai[RA], ak[RA], bk[RB], bj[RB];
        int64_t aoff[RA], boff[RB];  // offset of the element at k = 0 of the item
        for(int r = 0; r < RA; ++r) {
            const int idx = tid + NT * r;
            ai[r] = g.ta ? idx / BK : idx % BM;
            ak[r] = g.ta ? idx % BK : idx / BM;
            aoff[r] = g.ta ? int64_t(tl.m0 + ai[r]) * it.k + ak[r] : int64_t(ak[r]) * it.m + tl.m0 + ai[r];
        }
        for(int r = 0; r < RB; ++r) {
            const int idx = tid + NT * r;
            bj[r] = g.tb ? idx % BN : idx / BK;
            bk[r] = g.tb ? idx / BN : idx % BK;
            boff[r] = g.tb ? int64_t(bk[r]) * it.n + tl.n0 + bj[r] : int64_t(tl.n0 + bj[r]) * it.k + bk[r];
        }
        const int64_t astep = g.ta ? 1 : it.m, bstep = g.tb ? it.n : 1;  // offset per unit of k
        const double* A = g.A + it.a_off;
        const double* B = g.B + it.b_off;
        double ra[RA], rb[RB];
        auto fetch = [&](int k0) {
            if(full_mn && k0 + BK <= kend) {
                for(int r = 0; r < RA; ++r) ra[r] = A[aoff[r] + int64_t(k0) * astep];
                for(int r = 0; r < RB; ++r) rb[r] = B[boff[r] + int64_t(k0) * bstep];
            } else {
                for(int r = 0; r < RA; ++r)
                    ra[r] = (tl.m0 + ai[r] < it.m && k0 + ak[r] < kend) ? A[aoff[r] + int64_t(k0) * astep] : 0.0;
                for(int r = 0; r < RB; ++r)
                    rb[r] = (tl.n0 + bj[r] < it.n && k0 + bk[r] < kend) ? B[boff[r] + int64_t(k0) * bstep] : 0.0;
            }
        };
        auto put = [&](int s) {
            for(int r = 0; r < RA; ++r) sA[s][ak[r]][ai[r]] = ra[r];
            for(int r = 0; r < RB; ++r) sB[s][bk[r]][bj[r]] = rb[r];
        };
        d4 acc[MI][NI];
        for(int a = 0; a < MI; ++a)
            for(int b = 0; b < NI; ++b) acc[a][b] = d4{0, 0, 0, 0};
        __syncthreads();  // the previous tile's last reads of stage 0 are done
        if(kbeg < kend) {
            fetch(kbeg);
            put(0);
        }
        __syncthreads();
        int cur = 0;
        for(int k0 = kbeg; k0 < kend; k0 += BK) {
            const bool more = k0 + BK < kend;
            if(more) fetch(k0 + BK);
            for(int kq = 0; kq < BK; kq += 4) {
                const int kr = kq + (lane >> 4);
                double av[MI], bv[NI];
                for(int mi = 0; mi < MI; ++mi) av[mi] = sA[cur][kr][wm * WTM + mi * 16 + (lane & 15)];
                for(int ni = 0; ni < NI; ++ni) bv[ni] = sB[cur][kr][wn * WTN + ni * 16 + (lane & 15)];
                for(int mi = 0; mi < MI; ++mi)
                    for(int ni = 0; ni < NI; ++ni) acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
            }
            if(more) put(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
        const bool part = ex.split > 1;
        double* base = part ? g.ws + ex.ws_off + int64_t(tl.ks) * it.m * it.n : g.C + it.c_off;
        const double alpha = part ? 1.0 : g.alpha, beta = part ? 0.0 : g.beta;
        for(int mi = 0; mi < MI; ++mi)
            for(int ni = 0; ni < NI; ++ni)
                for(int r = 0; r < 4; ++r) {
                    const int row = tl.m0 + wm * WTM + mi * 16 + (lane >> 4) + 4 * r;
                    const int col = tl.n0 + wn * WTN + ni * 16 + (lane & 15);
                    if(row < it.m && col < it.n) {
                        double* p = base + int64_t(col) * it.m + row;
                        const double v = acc[mi][ni][r];
                        *p = (beta == 0.0) ? alpha * v : alpha * v + beta * *p;
                    }
                }
    }
}

// ---------------------------------------------------------------- FP64, skinny items
// Items whose output has one short side (S = min(m, n) <= SK_MAX) and a short K (<= SK_KMAX) are
// bound by the bytes of the long operand and of C, not by arithmetic: a 16 x 16 matrix-core tile
// would waste most of its work on them.  One thread per index of the long side (256 per
// workgroup) reads that index's K entries of the long operand once, keeps the S sums in registers
// with the small operand staged in LDS (read as broadcasts), and writes its S outputs.
constexpr int SK_MAX = 32, SK_KMAX = 32, SK_THREADS = 256;

// The S <= MS sums of one long-side index against the small operand in LDS (sS, [kk][s]) and
// hands output s to store(s, sum), in passes of at most 16 outputs: the K loaded
// elements stay in registers and only 16 accumulators are live (the gather kernel: 145 instead of
// 169 VGPRs at MS = 32, three waves per SIMD instead of two).  Every sum runs over kk in order, as
// one pass would.
template <int MS, class Store>
__device__ __forceinline__ void skinny_sums(const double* __restrict__ sS, const double (&x)[SK_KMAX], int K, int S, Store&& store)
{
    constexpr int MH = MS > 16 ? 16 : MS;
#pragma unroll
    for(int h = 0; h < MS; h += MH) {
        if(h >= S) break;
        double acc[MH];
#pragma unroll
        for(int c = 0; c < MH; ++c) acc[c] = 0.0;
#pragma unroll
        for(int kk = 0; kk < SK_KMAX; ++kk)
            if(kk < K)
#pragma unroll
                for(int c = 0; c < MH; ++c) acc[c] = __builtin_fma(sS[kk * MS + h + c], x[kk], acc[c]);
#pragma unroll
        for(int c = 0; c < MH; ++c)
            if(h + c < S) store(h + c, acc[c]);
    }
}

template <int MS, int R>
__global__ __launch_bounds__(SK_THREADS) void k_gemm_skinny(GemmArgs g, const Tile* __restrict__ blocks, int nblocks)
{
    __shared__ double sS[SK_KMAX * MS];  // the small operand, [kk][s]
    for(int b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const Tile tl = blocks[b];
        const xh_gemm_item it = g.items[tl.item];
        const bool wide = it.m <= it.n;  // wide: A is the small operand (C has S rows); else B
        const int S = wide ? it.m : it.n, K = it.k, Lng = wide ? it.n : it.m;
        __syncthreads();  // the previous block's reads of sS are done
        for(int e = threadIdx.x; e < K * MS; e += SK_THREADS) {
            const int sx = e % MS, kk = e / MS;
            double v = 0.0;
            if(sx < S) {
                if(wide)
                    v = g.ta ? g.A[it.a_off + int64_t(sx) * K + kk] : g.A[it.a_off + int64_t(kk) * it.m + sx];
                else
                    v = g.tb ? g.B[it.b_off + int64_t(kk) * it.n + sx] : g.B[it.b_off + int64_t(sx) * K + kk];
            }
            sS[kk * MS + sx] = v;
        }
        __syncthreads();
        // R long indices per thread, SK_THREADS apart, so that consecutive lanes stay consecutive; all
        // K loads of a thread are issued before the first product, to keep many loads in flight
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int L = tl.m0 + threadIdx.x + r * SK_THREADS;
            if(L >= Lng) break;
            double x[SK_KMAX];
#pragma unroll
            for(int kk = 0; kk < SK_KMAX; ++kk) {
                x[kk] = 0.0;
                if(kk < K) {
                    if(wide)
                        x[kk] = g.tb ? g.B[it.b_off + int64_t(kk) * it.n + L] : g.B[it.b_off + int64_t(L) * K + kk];
                    else
                        x[kk] = g.ta ? g.A[it.a_off + int64_t(L) * K + kk] : g.A[it.a_off + int64_t(kk) * it.m + L];
                }
            }
            double acc[MS];
#pragma unroll
            for(int sx = 0; sx < MS; ++sx) acc[sx] = 0.0;
#pragma unroll
            for(int kk = 0; kk < SK_KMAX; ++kk)
                if(kk < K)
#pragma unroll
                    for(int sx = 0; sx < MS; ++sx) acc[sx] = __builtin_fma(sS[kk * MS + sx], x[kk], acc[sx]);
#pragma unroll
            for(int sx = 0; sx < MS; ++sx)
                if(sx < S) {
                    double* p = g.C + it.c_off + (wide ? int64_t(L) * it.m + sx : int64_t(sx) * it.m + L);
                    *p = (g.beta == 0.0) ? g.alpha * acc[sx] : g.alpha * acc[sx] + g.beta * *p;
                }
        }
    }
}

// long-side indices per thread of the skinny kernel: one.  (Two or four, processed one after the
// other, halved the workgroups without adding loads in flight: on single items of 262144 x 20 x 20
// and 65536 x 20 x 20 one index per thread took 28.8 and 13.3 us against 30.5 and 21.0 us.)
constexpr int sk_rows(int) { return 1; }

// ---------------------------------------------------------------- FP64, skinny items, A through a permute
// The skinny kernel with its long operand A read through a single-source permute of src, instead
// of from the materialised permuted tensor (xh_gemm_gather).  Row r of an item's A block lies in
// domain tree a and column kk in codomain tree b; the element is entry (r - start_a, kk - start_b)
// of permute task task_base + a * ncod + b, and its source offset is
//   src_off(t) + sum over legs of (leg index) * (leg stride in the source, times src_ld for a
//   source column leg),
// the domain legs' indices peeled off the row, the codomain legs' off the column.  The plan cuts
// every item's rows into blocks of at most 256 rows that meet at most GA_SEG domain trees.  Per
// (block tree, column) the block tabulates in LDS everything but the domain legs' terms (the
// column's offset) and the domain legs' strides; a thread then adds (index x stride) over its
// row's domain legs per column.  Coefficients are +1 or -1, so the operand is the source element,
// negated for -1: the value the permute writes, up to the sign of a zero, which sums that start
// from +0 do not see.  The sums run in the skinny kernel's order, so the product equals
// permute + xh_gemm bit for bit.
constexpr int GA_MAXR = 8;  // domain legs of a gather plan's permuted tensor (checked at plan creation)
constexpr int GA_SEG = 4;   // domain trees one block may meet
struct GatherBlock {
    int32_t item, row0, nrows, nseg;
    int32_t a[GA_SEG];      // domain tree of each segment
    int32_t start[GA_SEG];  // first row of that tree (segments ascend)
};
struct GatherItem {
    int32_t task_base, ncod;
    int64_t cols_off;  // per column kk < K: codomain tree b | (column inside it << 16)
};
struct GatherSrc {
    int64_t off, ld;   // source subblock offset and leading dimension
    int32_t neg, pad;  // 1: coefficient -1
};
struct GatherDev {
    const GatherBlock* blocks;
    const GatherItem* items;  // per GEMM item
    const int32_t* cols;
    const xh::LegInfo* legs;  // ntasks * nlegs
    const GatherSrc* src;     // per task
    int32_t nlegs, new_rank, nblocks;
};

// The permute of the product folded into the stores (SCATTER): every product subblock (domain tree
// a, codomain tree c) is the single source of one task of that permute, which fixes its place in
// the output.  Per (block tree, product column) the block tabulates the column's destination
// offset and the destination strides of the product's domain legs, which are the permuted left
// operand's domain legs, so the row's indices are the ones decoded for the gather.
struct ScatterItem {
    int64_t grid_off, cols_off;  // product subblock (a, c) -> task grid[grid_off + a * ncod + c]; columns as GatherItem's
    int32_t ncod, pad;
};
struct ScatterDev {
    const ScatterItem* items;  // per GEMM item
    const int32_t* grid;
    const int32_t* cols;       // per product column: codomain tree | (column inside it << 16)
    const xh::LegInfo* legs;   // per task and product leg j: extent, division and destination stride of leg j
    const int64_t* dst_off;    // per task: offset of its output subblock
    const int32_t* neg;        // per task: 1 for coefficient -1
    int32_t nlegs;             // product legs (its domain legs first)
};

template <int MS, bool SCATTER>
__global__ __launch_bounds__(SK_THREADS, SCATTER ? 1 : 4) void k_gemm_skinny_gather(GemmArgs g, GatherDev gd, const double* __restrict__ src, ScatterDev sd)
{
    constexpr int DS = SCATTER ? GA_SEG : 1, DC = SCATTER ? SK_MAX : 1, DR = SCATTER ? GA_MAXR : 1;
    __shared__ int64_t s_doff[DS][DC];      // per tree and product column: destination offset of the column
    __shared__ int64_t s_drs[DS][DC][DR];   // per tree and product column: destination strides of the domain legs
    __shared__ uint32_t s_dneg[DS];         // bit c: the column's task has coefficient -1
    __shared__ double sS[SK_KMAX * MS];                     // B, [kk][s]
    __shared__ int64_t s_off[GA_SEG][SK_KMAX];              // per tree and column: the column's offset
    __shared__ int64_t s_rs[GA_SEG][SK_KMAX][GA_MAXR];      // per tree and column: the domain legs' strides
    __shared__ xh::LegInfo s_rleg[GA_SEG][GA_MAXR];         // per tree: its domain legs (extents, divisions)
    __shared__ uint32_t s_neg[GA_SEG];                      // bit kk: the column's task has coefficient -1
    const int R = gd.new_rank, NL = gd.nlegs;
    for(int blk = blockIdx.x; blk < gd.nblocks; blk += gridDim.x) {
        const GatherBlock bl = gd.blocks[blk];
        const xh_gemm_item it = g.items[bl.item];
        const GatherItem gi = gd.items[bl.item];
        const int S = it.n, K = it.k;
        __syncthreads();  // the previous block's reads of the shared tables are done
        for(int e = threadIdx.x; e < K * MS; e += SK_THREADS) {
            const int sx = e % MS, kk = e / MS;
            double v = 0.0;
            if(sx < S) v = g.tb ? g.B[it.b_off + int64_t(kk) * it.n + sx] : g.B[it.b_off + int64_t(sx) * K + kk];
            sS[kk * MS + sx] = v;
        }
        if(int(threadIdx.x) < GA_SEG) s_neg[threadIdx.x] = 0;
        if constexpr(SCATTER)
            if(int(threadIdx.x) < GA_SEG) s_dneg[threadIdx.x] = 0;
        __syncthreads();
        for(int e = threadIdx.x; e < bl.nseg * K; e += SK_THREADS) {
            const int sg = e / K, kk = e % K;
            const int32_t cb = gd.cols[gi.cols_off + kk];
            const int t = gi.task_base + gd.blocks[blk].a[sg] * gi.ncod + (cb & 0xffff);
            const GatherSrc so = gd.src[t];
            const xh::LegInfo* Lg = gd.legs + int64_t(t) * NL;
            // the codomain legs' terms of column kk (its index inside tree b peeled leg by leg)
            uint32_t x = static_cast<uint32_t>(cb >> 16);
            int64_t off = so.off;
            for(int k = R; k < NL; ++k) {
                const xh::LegInfo li = Lg[k];
                const uint32_t q = xh::fdiv(x, li.magic, xh::leg_shift(li));
                const uint32_t j = x - q * li.ext;
                x = q;
                off += (li.flags & 256u) ? int64_t(j) * li.stride * so.ld : int64_t(j) * li.stride;
            }
            s_off[sg][kk] = off;
            for(int k = 0; k < R; ++k) {
                const xh::LegInfo li = Lg[k];
                s_rs[sg][kk][k] = (li.flags & 256u) ? int64_t(li.stride) * so.ld : int64_t(li.stride);
                if(kk == 0) s_rleg[sg][k] = li;
            }
            if(so.neg) atomicOr(&s_neg[sg], 1u << kk);
        }
        if constexpr(SCATTER) {
            const ScatterItem si = sd.items[bl.item];
            const int NP = sd.nlegs;
            for(int e = threadIdx.x; e < bl.nseg * S; e += SK_THREADS) {
                const int sg = e / S, c = e % S;
                const int32_t cb = sd.cols[si.cols_off + c];
                const int t = sd.grid[si.grid_off + int64_t(gd.blocks[blk].a[sg]) * si.ncod + (cb & 0xffff)];
                const xh::LegInfo* Lg = sd.legs + int64_t(t) * NP;
                uint32_t x = static_cast<uint32_t>(cb >> 16);
                int64_t off = sd.dst_off[t];
                for(int k = R; k < NP; ++k) {
                    const xh::LegInfo li = Lg[k];
                    const uint32_t q = xh::fdiv(x, li.magic, xh::leg_shift(li));
                    off += int64_t(x - q * li.ext) * li.stride;
                    x = q;
                }
                s_doff[sg][c] = off;
                for(int k = 0; k < R; ++k) s_drs[sg][c][k] = Lg[k].stride;
                if(sd.neg[t]) atomicOr(&s_dneg[sg], 1u << c);
            }
        }
        __syncthreads();
        if(int(threadIdx.x) >= bl.nrows) continue;  // no barrier follows before the next block's first one
        const int L = bl.row0 + threadIdx.x;
        int sg = 0, first = bl.start[0];
#pragma unroll
        for(int q = 1; q < GA_SEG; ++q)
            if(q < bl.nseg && bl.start[q] <= L) {
                sg = q;
                first = bl.start[q];
            }
        // the row's domain-leg indices inside its tree
        uint32_t x = static_cast<uint32_t>(L - first);
        uint32_t dig[GA_MAXR];
#pragma unroll
        for(int k = 0; k < GA_MAXR; ++k) {
            dig[k] = 0;
            if(k < R) {
                const xh::LegInfo li = s_rleg[sg][k];
                const uint32_t q = xh::fdiv(x, li.magic, xh::leg_shift(li));
                dig[k] = x - q * li.ext;
                x = q;
            }
        }
        double xv[SK_KMAX];
#pragma unroll
        for(int kk = 0; kk < SK_KMAX; ++kk) {
            xv[kk] = 0.0;
            if(kk < K) {
                int64_t off = s_off[sg][kk];
#pragma unroll
                for(int k = 0; k < GA_MAXR; ++k)
                    if(k < R) off += int64_t(dig[k]) * s_rs[sg][kk][k];
                xv[kk] = src[off];
            }
        }
        const uint32_t negm = s_neg[sg];
#pragma unroll
        for(int kk = 0; kk < SK_KMAX; ++kk)
            if((negm >> kk) & 1u) xv[kk] = -xv[kk];
        if constexpr(!SCATTER) {
            double* base = g.C + it.c_off + L;
            const int64_t ld = it.m;
            skinny_sums<MS>(sS, xv, K, S, [&](int c, double a) {
                double* p = base + int64_t(c) * ld;
                *p = (g.beta == 0.0) ? g.alpha * a : g.alpha * a + g.beta * *p;
            });
        } else {
            // each output straight to its place in the permuted product: the column's offset plus
            // (domain-leg index x destination stride) over the row's legs; 0 + c v with c = +-1,
            // as the permute writes it (beta is 0)
            const uint32_t dneg = s_dneg[sg];
            skinny_sums<MS>(sS, xv, K, S, [&](int c, double a) {
                int64_t off = s_doff[sg][c];
#pragma unroll
                for(int k = 0; k < GA_MAXR; ++k)
                    if(k < R) off += int64_t(dig[k]) * s_drs[sg][c][k];
                const double v = g.alpha * a;
                g.C[off] = ((dneg >> c) & 1u) ? 0.0 - v : 0.0 + v;
            });
        }
    }
}

// ---------------------------------------------------------------- FP64, 128 x 128 tiles
// For items with both output sides >= 192: 8 waves in 2 x 4, each owning 64 x 32 (4 x 2 MFMA
// blocks, 8 accumulators, so two or three waves share a SIMD), K steps of 16 staged in LDS, and
// the next K step's operands fetched into registers while the matrix cores work on the current
// one.  (A 4-wave variant with 64 x 64 per wave ran at one wave per SIMD and was 25-30 % slower.)
constexpr int BM2 = 128, BN2 = 128;
constexpr int BIG_MIN = 192;

constexpr int G8_THREADS = 512;
__global__ __launch_bounds__(G8_THREADS, 4) void k_gemm_f64_128(GemmArgs g)
{
    constexpr int BK = 16;
    __shared__ double sA[BK][BM2 + 1];
    __shared__ double sB[BK][BN2 + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;  // wn in 0..3
    const TileRange tr = my_tiles(g);
    for(int t = tr.begin; t < tr.end; t += tr.stride) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        const ItemExt ex = g.ext[tl.item];
        const int kbeg = tl.ks * ex.kc, kend = min(it.k, kbeg + ex.kc);
        d4 acc[4][2];
        for(int a = 0; a < 4; ++a)
            for(int b = 0; b < 2; ++b) acc[a][b] = d4{0, 0, 0, 0};
        double ra[4], rb[4];
        auto fetch = [&](int k0) {
            for(int r = 0; r < 4; ++r) {
                const int idx = tid + G8_THREADS * r;
                const int i = g.ta ? idx >> 4 : idx & 127, kk = g.ta ? idx & 15 : idx >> 7;
                ra[r] = load_a(g, it, tl.m0 + i, k0 + kk, kend);
                const int kb = g.tb ? idx >> 7 : idx & 15, j = g.tb ? idx & 127 : idx >> 4;
                rb[r] = load_b(g, it, k0 + kb, tl.n0 + j, kend);
            }
        };
        if(kbeg < kend) fetch(kbeg);
        for(int k0 = kbeg; k0 < kend; k0 += BK) {
            __syncthreads();
            for(int r = 0; r < 4; ++r) {
                const int idx = tid + G8_THREADS * r;
                const int i = g.ta ? idx >> 4 : idx & 127, kk = g.ta ? idx & 15 : idx >> 7;
                sA[kk][i] = ra[r];
                const int kb = g.tb ? idx >> 7 : idx & 15, j = g.tb ? idx & 127 : idx >> 4;
                sB[kb][j] = rb[r];
            }
            __syncthreads();
            if(k0 + BK < kend) fetch(k0 + BK);
            for(int kq = 0; kq < BK; kq += 4) {
                const int kr = kq + (lane >> 4);
                double av[4], bv[2];
                for(int mi = 0; mi < 4; ++mi) av[mi] = sA[kr][wm * 64 + mi * 16 + (lane & 15)];
                for(int ni = 0; ni < 2; ++ni) bv[ni] = sB[kr][wn * 32 + ni * 16 + (lane & 15)];
                for(int mi = 0; mi < 4; ++mi)
                    for(int ni = 0; ni < 2; ++ni) acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
            }
        }
        const bool part = ex.split > 1;
        double* base = part ? g.ws + ex.ws_off + int64_t(tl.ks) * it.m * it.n : g.C + it.c_off;
        const double alpha = part ? 1.0 : g.alpha, beta = part ? 0.0 : g.beta;
        for(int mi = 0; mi < 4; ++mi)
            for(int ni = 0; ni < 2; ++ni)
                for(int r = 0; r < 4; ++r) {
                    const int row = tl.m0 + wm * 64 + mi * 16 + (lane >> 4) + 4 * r;
                    const int col = tl.n0 + wn * 32 + ni * 16 + (lane & 15);
                    if(row < it.m && col < it.n) {
                        double* p = base + int64_t(col) * it.m + row;
                        const double v = acc[mi][ni][r];
                        *p = (beta == 0.0) ? alpha * v : alpha * v + beta * *p;
                    }
                }
    }
}

// ---------------------------------------------------------------- FP64 through int8 slices
// Ozaki, Ogita, Oishi and Rump, Numer. Algorithms 59 (2012) 95-118 (error-free splitting of a
// matrix product), with int8 slices as in Ootomo, Ozaki and Yokota, IJHPCA (2024).  Row i of A is
// 2^ea_i sum_s a_s,i 2^-7(s+1) with |a_s| <= 127 (s = 0..OZ_NS-1), column j of B likewise; every
// slice product is an exact integer matrix, and the products of one level d = s + t (weight
// 2^-7(d+2)) are summed in int32, exact while OZ_NS * OZ_FLUSH * OZ_KB * 127^2 < 2^31; every
// OZ_FLUSH K steps the levels are folded into FP64 accumulators, smallest weight first.  The
// levels d >= OZ_NS and the slicing remainder are dropped: both are below k 2^-7 OZ_NS max|A_i.|
// max|B_.j| (2^-56 per term), under FP64 GEMM's own k 2^-53 bound.
constexpr int OZ_NS = 8;
constexpr int OZ_KB = 64;            // the K of one v_mfma_i32_16x16x64_i8
constexpr int OZ_PITCH = OZ_KB + 16;  // LDS row pitch in bytes
constexpr int OZ_FLUSH = 64;          // K steps between int32 -> FP64 folds
constexpr int32_t OZ_NONFINITE = -0x40000000;
static_assert(double(OZ_NS) * OZ_FLUSH * OZ_KB * 127 * 127 < 2147483648.0, "int32 level sums must stay exact");

struct OzItem {
    int64_t sa_off, sb_off;  // byte offsets of the item's A and B slices (OZ_NS x rows x kp)
    int64_t ea_off, eb_off;  // offsets of its row and column exponents
    int32_t kp, pad;         // K rounded up to OZ_KB (zero padded)
};
// one slicing job: rows [row0, row0 + nrows) of operand op (0: rows of A, 1: columns of B) of an
// item; g = lanes per row for operands whose K direction is contiguous in memory
struct OzJob {
    int32_t item, op, row0, g;
};
struct OzArgs {
    const OzItem* oz;
    int8_t* sl;
    int32_t* ex;
};
typedef int i4v __attribute__((ext_vector_type(4)));

// the row's base pointer and K stride, and the number of rows, of operand op of an item
__device__ __forceinline__ const double* oz_row(const GemmArgs& g, const xh_gemm_item& it, int op, int r, int64_t& step)
{
    if(op == 0) {
        step = g.ta ? 1 : it.m;
        return g.ta ? g.A + it.a_off + int64_t(r) * it.k : g.A + it.a_off + r;
    }
    step = g.tb ? it.n : 1;
    return g.tb ? g.B + it.b_off + r : g.B + it.b_off + int64_t(r) * it.k;
}

__device__ __forceinline__ int oz_exponent(double mx, int bad)
{
    int e = 0;
    if(bad) return OZ_NONFINITE;
    if(mx > 0.0) frexp(mx, &e);  // mx = f 2^e with f in [1/2, 1): every |x| / 2^e < 1
    return e;
}

// 8 slices of x (|x| < 1): x = sum_s q_s 2^-7(s+1) + remainder, |q_s| <= 127
__device__ __forceinline__ void oz_cut(double x, int8_t (&q)[OZ_NS])
{
#pragma unroll
    for(int s = 0; s < OZ_NS; ++s) {
        x *= 128.0;  // exact
        const double t = trunc(x);
        x -= t;  // exact: |x - t| < 1 and both carry at most 53 bits
        q[s] = static_cast<int8_t>(static_cast<int>(t));
    }
}

// operands whose rows are contiguous (A not transposed, B transposed): one thread per row, K
// walked with coalesced loads across the workgroup's rows; slices written 4 K entries at a time
__global__ __launch_bounds__(256) void k_oz_split_rows(GemmArgs g, OzArgs o, const OzJob* __restrict__ jobs)
{
    const OzJob jb = jobs[blockIdx.x];
    const xh_gemm_item it = g.items[jb.item];
    const OzItem oi = o.oz[jb.item];
    const int rows = jb.op == 0 ? it.m : it.n;
    const int r = jb.row0 + int(threadIdx.x);
    if(r >= rows) return;
    int64_t step;
    const double* src = oz_row(g, it, jb.op, r, step);
    double mx = 0.0;
    int bad = 0;
    for(int kk = 0; kk < it.k; ++kk) {
        const double v = src[kk * step];
        bad |= !isfinite(v);
        mx = fmax(mx, fabs(v));
    }
    const int e = oz_exponent(mx, bad);
    o.ex[(jb.op == 0 ? oi.ea_off : oi.eb_off) + r] = e;
    const int es = bad ? 0 : e;
    int8_t* out = o.sl + (jb.op == 0 ? oi.sa_off : oi.sb_off) + int64_t(r) * oi.kp;
    const int64_t sstride = int64_t(rows) * oi.kp;
    for(int kk = 0; kk < oi.kp; kk += 4) {
        int8_t q[4][OZ_NS];
#pragma unroll
        for(int u = 0; u < 4; ++u) oz_cut((kk + u < it.k && !bad) ? ldexp(src[(kk + u) * step], -es) : 0.0, q[u]);
#pragma unroll
        for(int s = 0; s < OZ_NS; ++s) {
            const uint32_t w = uint32_t(uint8_t(q[0][s])) | (uint32_t(uint8_t(q[1][s])) << 8) | (uint32_t(uint8_t(q[2][s])) << 16) |
                               (uint32_t(uint8_t(q[3][s])) << 24);
            *reinterpret_cast<uint32_t*>(out + s * sstride + kk) = w;
        }
    }
}

// operands whose K direction is contiguous (A transposed, B not): a group of g lanes per row
__global__ __launch_bounds__(256) void k_oz_split_k(GemmArgs g, OzArgs o, const OzJob* __restrict__ jobs)
{
    const OzJob jb = jobs[blockIdx.x];
    const xh_gemm_item it = g.items[jb.item];
    const OzItem oi = o.oz[jb.item];
    const int rows = jb.op == 0 ? it.m : it.n;
    const int gs = jb.g, lig = int(threadIdx.x) % gs;
    const int r = jb.row0 + int(threadIdx.x) / gs;
    const bool live = r < rows;
    int64_t step = 0;
    const double* src = live ? oz_row(g, it, jb.op, r, step) : nullptr;
    double mx = 0.0;
    int bad = 0;
    if(live)
        for(int kk = lig; kk < it.k; kk += gs) {
            const double v = src[kk];
            bad |= !isfinite(v);
            mx = fmax(mx, fabs(v));
        }
    // reduce within the lane group (gs divides 64: a group never spans two waves)
    for(int off = gs >> 1; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off));
        bad |= __shfl_xor(bad, off);
    }
    if(!live) return;
    const int e = oz_exponent(mx, bad);
    if(lig == 0) o.ex[(jb.op == 0 ? oi.ea_off : oi.eb_off) + r] = e;
    const int es = bad ? 0 : e;
    int8_t* out = o.sl + (jb.op == 0 ? oi.sa_off : oi.sb_off) + int64_t(r) * oi.kp;
    const int64_t sstride = int64_t(rows) * oi.kp;
    for(int kk = lig; kk < oi.kp; kk += gs) {
        int8_t q[OZ_NS];
        oz_cut((kk < it.k && !bad) ? ldexp(src[kk], -es) : 0.0, q);
#pragma unroll
        for(int s = 0; s < OZ_NS; ++s) out[s * sstride + kk] = q[s];
    }
}


constexpr int OZ_THREADS = 512;                       // 8 waves: 2 x 4 wave tiles of 32 x 16
constexpr int OZ_UNITS = OZ_NS * 64 * 4 / OZ_THREADS;  // 16-byte units per thread and operand per K step

// the 16-byte units of one K step that thread tid stages: unit u = tid + OZ_THREADS v is slice
// u / 256, tile row (u / 4) % 64, bytes 16 (u % 4) .. of the step; bytes outside [kbeg, kend)
// (a K chunk's ends) are zeroed
__device__ __forceinline__ void oz_fetch(const OzArgs& o, const OzItem& oi, const xh_gemm_item& it, const Tile& tl, int k0, int kbeg, int kend,
                                         i4v (&va)[OZ_UNITS], i4v (&vb)[OZ_UNITS])
{
#pragma unroll
    for(int v = 0; v < OZ_UNITS; ++v) {
        const int u = int(threadIdx.x) + OZ_THREADS * v;
        const int sl = u >> 8, row = (u >> 2) & 63, c = u & 3;
        const int i = tl.m0 + row, j = tl.n0 + row, kk0 = k0 + 16 * c;
        va[v] = i4v{0, 0, 0, 0};
        vb[v] = i4v{0, 0, 0, 0};
        if(i < it.m) va[v] = *reinterpret_cast<const i4v*>(o.sl + oi.sa_off + (int64_t(sl) * it.m + i) * oi.kp + kk0);
        if(j < it.n) vb[v] = *reinterpret_cast<const i4v*>(o.sl + oi.sb_off + (int64_t(sl) * it.n + j) * oi.kp + kk0);
        if(kk0 < kbeg || kk0 + 16 > kend) {
            const int lo = kbeg - kk0, hi = kend - kk0;  // keep bytes b with lo <= b < hi
#pragma unroll
            for(int w = 0; w < 4; ++w) {
                uint32_t m = 0;
#pragma unroll
                for(int b = 0; b < 4; ++b)
                    if(4 * w + b >= lo && 4 * w + b < hi) m |= 0xffu << (8 * b);
                va[v][w] = int(uint32_t(va[v][w]) & m);
                vb[v][w] = int(uint32_t(vb[v][w]) & m);
            }
        }
    }
}

__device__ __forceinline__ void oz_fold(i4v (&acc)[OZ_NS][2], d4 (&facc)[2])
{
#pragma unroll
    for(int d = OZ_NS - 1; d >= 0; --d) {
        const double w = ldexp(1.0, -7 * (d + 2));
#pragma unroll
        for(int mi = 0; mi < 2; ++mi) {
#pragma unroll
            for(int r = 0; r < 4; ++r) facc[mi][r] += w * static_cast<double>(acc[d][mi][r]);
            acc[d][mi] = i4v{0, 0, 0, 0};
        }
    }
}

// 64 x 64 output tiles, 8 waves in 2 x 4, each 2 MFMA tiles of 16 x 16 (32 x 16); K steps of 64 with
// all OZ_NS slices of both operands staged in LDS (80 KiB); the next step's slices are fetched
// into registers while the matrix cores work on the current one
__global__ __launch_bounds__(OZ_THREADS) void k_gemm_oz(GemmArgs g, OzArgs o)
{
    static_assert(OZ_NS * 64 * 4 == OZ_UNITS * OZ_THREADS, "a K step is OZ_UNITS 16-byte units per thread and operand");
    __shared__ __attribute__((aligned(16))) int8_t sA[OZ_NS][BM][OZ_PITCH];
    __shared__ __attribute__((aligned(16))) int8_t sB[OZ_NS][BN][OZ_PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const TileRange tr = my_tiles(g);
    for(int t = tr.begin; t < tr.end; t += tr.stride) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        const ItemExt ex = g.ext[tl.item];
        const OzItem oi = o.oz[tl.item];
        const int kbeg = tl.ks * ex.kc, kend = min(it.k, kbeg + ex.kc);
        d4 facc[2];
        i4v acc[OZ_NS][2];
#pragma unroll
        for(int mi = 0; mi < 2; ++mi) {
            facc[mi] = d4{0, 0, 0, 0};
#pragma unroll
            for(int d = 0; d < OZ_NS; ++d) acc[d][mi] = i4v{0, 0, 0, 0};
        }
        int nsteps = 0;
        const int kfirst = kbeg - (kbeg % OZ_KB);
        i4v va[OZ_UNITS], vb[OZ_UNITS];
        if(kfirst < kend) oz_fetch(o, oi, it, tl, kfirst, kbeg, kend, va, vb);
        for(int k0 = kfirst; k0 < kend; k0 += OZ_KB) {
#pragma unroll
            for(int v = 0; v < OZ_UNITS; ++v) {
                const int u = tid + OZ_THREADS * v;
                const int sl = u >> 8, row = (u >> 2) & 63, c = u & 3;
                *reinterpret_cast<i4v*>(&sA[sl][row][16 * c]) = va[v];
                *reinterpret_cast<i4v*>(&sB[sl][row][16 * c]) = vb[v];
            }
            __syncthreads();
            if(k0 + OZ_KB < kend) oz_fetch(o, oi, it, tl, k0 + OZ_KB, kbeg, kend, va, vb);
            i4v bf[OZ_NS];
#pragma unroll
            for(int tt = 0; tt < OZ_NS; ++tt) bf[tt] = *reinterpret_cast<const i4v*>(&sB[tt][wn * 16 + (lane & 15)][16 * (lane >> 4)]);
#pragma unroll
            for(int sa = 0; sa < OZ_NS; ++sa) {
                i4v af[2];
#pragma unroll
                for(int mi = 0; mi < 2; ++mi) af[mi] = *reinterpret_cast<const i4v*>(&sA[sa][wm * 32 + mi * 16 + (lane & 15)][16 * (lane >> 4)]);
#pragma unroll
                for(int tt = 0; tt + sa < OZ_NS; ++tt)
#pragma unroll
                    for(int mi = 0; mi < 2; ++mi) acc[sa + tt][mi] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[mi], bf[tt], acc[sa + tt][mi], 0, 0, 0);
            }
            __syncthreads();
            if(++nsteps == OZ_FLUSH) {
                oz_fold(acc, facc);
                nsteps = 0;
            }
        }
        oz_fold(acc, facc);
        // scale by 2^(ea_i + eb_j) and store; register r of lane l holds row 4 (l / 16) + r,
        // column l % 16 of its 16 x 16 tile
        const bool part = ex.split > 1;
        double* base = part ? g.ws + ex.ws_off + int64_t(tl.ks) * it.m * it.n : g.C + it.c_off;
        const double alpha = part ? 1.0 : g.alpha, beta = part ? 0.0 : g.beta;
        const int col = tl.n0 + wn * 16 + (lane & 15);
        const int eb = col < it.n ? o.ex[oi.eb_off + col] : 0;
#pragma unroll
        for(int mi = 0; mi < 2; ++mi)
#pragma unroll
            for(int r = 0; r < 4; ++r) {
                const int row = tl.m0 + wm * 32 + mi * 16 + 4 * (lane >> 4) + r;
                if(row < it.m && col < it.n) {
                    const int ea = o.ex[oi.ea_off + row];
                    const double v = (ea == OZ_NONFINITE || eb == OZ_NONFINITE) ? __builtin_nan("") : ldexp(facc[mi][r], ea + eb);
                    double* pc = base + int64_t(col) * it.m + row;
                    *pc = (beta == 0.0) ? alpha * v : alpha * v + beta * *pc;
                }
            }
    }
}

// ---------------------------------------------------------------- FP32
__global__ __launch_bounds__(G_THREADS) void k_gemm_f32(GemmArgs g)
{
    constexpr int BK = 32;
    __shared__ float sA[BK][BM + 1];
    __shared__ float sB[BK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const TileRange tr = my_tiles(g);
    for(int t = tr.begin; t < tr.end; t += tr.stride) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        const ItemExt ex = g.ext[tl.item];
        const int kbeg = tl.ks * ex.kc, kend = min(it.k, kbeg + ex.kc);
        f4 acc[2][2];
        for(int a = 0; a < 2; ++a)
            for(int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
        for(int k0 = kbeg; k0 < kend; k0 += BK) {
            for(int r = 0; r < 8; ++r) {
                const int idx = tid + G_THREADS * r;
                const int i = g.ta ? idx >> 5 : idx & 63, kk = g.ta ? idx & 31 : idx >> 6;
                sA[kk][i] = static_cast<float>(load_a(g, it, tl.m0 + i, k0 + kk, kend));
                const int kb = g.tb ? idx >> 6 : idx & 31, j = g.tb ? idx & 63 : idx >> 5;
                sB[kb][j] = static_cast<float>(load_b(g, it, k0 + kb, tl.n0 + j, kend));
            }
            __syncthreads();
            for(int kq = 0; kq < BK; kq += 4) {
                const int kr = kq + (lane >> 4);
                float av[2], bv[2];
                for(int mi = 0; mi < 2; ++mi) av[mi] = sA[kr][wm * 32 + mi * 16 + (lane & 15)];
                for(int ni = 0; ni < 2; ++ni) bv[ni] = sB[kr][wn * 32 + ni * 16 + (lane & 15)];
                for(int mi = 0; mi < 2; ++mi)
                    for(int ni = 0; ni < 2; ++ni) acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
            }
            __syncthreads();
        }
        store_tile<false>(g, it, ex, tl, wm, wn, lane, acc);
    }
}

// ---------------------------------------------------------------- BF16x3
__device__ __forceinline__ unsigned short bf16_rn(float x)
{
    // round-to-nearest-even truncation of an IEEE binary32 to bfloat16 (finite inputs)
    unsigned int u = __float_as_uint(x);
    unsigned int lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float(static_cast<unsigned int>(h) << 16); }

__global__ __launch_bounds__(G_THREADS) void k_gemm_bf16x3(GemmArgs g)
{
    constexpr int BK = 32;
    // [k][i] with k packed by 4 for the 16x16x16 operand layout
    __shared__ unsigned short sAh[BK][BM + 4], sAl[BK][BM + 4];
    __shared__ unsigned short sBh[BK][BN + 4], sBl[BK][BN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const TileRange tr = my_tiles(g);
    for(int t = tr.begin; t < tr.end; t += tr.stride) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        const ItemExt ex = g.ext[tl.item];
        const int kbeg = tl.ks * ex.kc, kend = min(it.k, kbeg + ex.kc);
        f4 acc[2][2];
        for(int a = 0; a < 2; ++a)
            for(int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
        for(int k0 = kbeg; k0 < kend; k0 += BK) {
            for(int r = 0; r < 8; ++r) {
                const int idx = tid + G_THREADS * r;
                const int i = g.ta ? idx >> 5 : idx & 63, kk = g.ta ? idx & 31 : idx >> 6;
                float a = static_cast<float>(load_a(g, it, tl.m0 + i, k0 + kk, kend));
                unsigned short ah = bf16_rn(a);
                sAh[kk][i] = ah;
                sAl[kk][i] = bf16_rn(a - bf16_to_f(ah));
                const int kb = g.tb ? idx >> 6 : idx & 31, j = g.tb ? idx & 63 : idx >> 5;
                float b = static_cast<float>(load_b(g, it, k0 + kb, tl.n0 + j, kend));
                unsigned short bh = bf16_rn(b);
                sBh[kb][j] = bh;
                sBl[kb][j] = bf16_rn(b - bf16_to_f(bh));
            }
            __syncthreads();
            for(int kq = 0; kq < BK; kq += 16) {
                const int kr = kq + 4 * (lane >> 4);
                s4 ah[2], al[2], bh[2], bl[2];
                for(int mi = 0; mi < 2; ++mi) {
                    const int i = wm * 32 + mi * 16 + (lane & 15);
                    for(int e = 0; e < 4; ++e) {
                        ah[mi][e] = static_cast<short>(sAh[kr + e][i]);
                        al[mi][e] = static_cast<short>(sAl[kr + e][i]);
                    }
                }
                for(int ni = 0; ni < 2; ++ni) {
                    const int j = wn * 32 + ni * 16 + (lane & 15);
                    for(int e = 0; e < 4; ++e) {
                        bh[ni][e] = static_cast<short>(sBh[kr + e][j]);
                        bl[ni][e] = static_cast<short>(sBl[kr + e][j]);
                    }
                }
                for(int mi = 0; mi < 2; ++mi)
                    for(int ni = 0; ni < 2; ++ni) {
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                    }
            }
            __syncthreads();
        }
        store_tile<false>(g, it, ex, tl, wm, wn, lane, acc);
    }
}

}  // namespace

// A work list of tiles cut into per-XCD parts.
struct TileList {
    Tile* d_tiles = nullptr;
    int32_t ntiles = 0;
    int32_t part[NPART + 1] = {};
};

struct xh_gemm_plan : xh::PlanLife {
    xh::PackedUpload up;
    xh_gemm_item* d_items = nullptr;
    ItemExt* d_ext = nullptr;
    RedItem* d_red = nullptr;
    TileList all64;           // every item in 64 x 64 tiles (FP32 / BF16x3 modes), built on first use
    TileList big128, rest64;  // FP64: items with both sides >= BIG_MIN in 128 x 128 tiles, the others in 64 x 64
    TileList skinny;          // FP64: blocks of 256 long-side indices of the skinny items (Tile.m0 = first index)
    int sk_ms = 8;            // register width of the skinny kernel: 8, 16 or 32 >= every skinny item's short side
    bool all64_built = false, any_tiles = false;
    xh::PackedUpload up_all;
    // XH_OZAKI, built on first use: the slice layout of the items that take the int8 path, their
    // tiles and slicing jobs, and the 64 x 64 FP64 tiles of the other non-skinny items
    bool oz_built = false;
    xh::PackedUpload up_oz;
    OzItem* d_oz = nullptr;
    TileList oz_tiles, oz_rest;
    OzJob* d_jobs_rows = nullptr;
    OzJob* d_jobs_k = nullptr;
    int32_t n_jobs_rows = 0, n_jobs_k = 0;
    int64_t oz_slice_bytes = 0, oz_exps = 0;
    std::vector<xh_gemm_item> h_items;  // host copies for the deferred list
    std::vector<ItemExt> h_ext;
    std::vector<int> h_order;
    int32_t nred = 0;
    int64_t ws_doubles = 0, max_red_mn = 0;
    int ta = 0, tb = 0;
    uint64_t sig = 0;  // autotuner key: transposes, split-K use and log2 buckets of the plan's sizes
    double flops = 0.0;
};

namespace {
// contiguous parts of equal work (sum of K over the tiles), one per XCD
void cut_parts(const std::vector<double>& work, int32_t part[NPART + 1])
{
    double total = 0.0;
    for(double w : work) total += w;
    part[0] = 0;
    double acc = 0.0;
    int x = 1;
    for(size_t t = 0; t < work.size() && x < NPART; ++t) {
        acc += work[t];
        while(x < NPART && acc >= total * x / NPART) part[x++] = static_cast<int32_t>(t + 1);
    }
    for(; x <= NPART; ++x) part[x] = static_cast<int32_t>(work.size());
}

bool is_big(const xh_gemm_item& it) { return it.m >= BIG_MIN && it.n >= BIG_MIN; }
bool is_skinny(const xh_gemm_item& it) { return std::min(it.m, it.n) <= SK_MAX && it.k <= SK_KMAX; }
}  // namespace

extern "C" int xh_gemm_plan_create(xh_ctx* ctx, const xh_gemm_item* items, int64_t nitems, int trans_a, int trans_b, xh_gemm_plan** out)
{
    *out = nullptr;
    auto* plan = new xh_gemm_plan();
    int rc = guard(ctx, [&] {
        xh::refuse_under_capture(ctx, "plan creation");
        std::vector<xh_gemm_item> v(items, items + nitems);
        // largest K*tiles first so each part's loop ends with small work
        std::vector<int> order(v.size());
        for(size_t i = 0; i < v.size(); ++i) order[i] = static_cast<int>(i);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return double(v[a].m) * v[a].n * v[a].k > double(v[b].m) * v[b].n * v[b].k; });
        int64_t base_tiles = 0;  // output tiles of the FP64 launch
        for(const auto& it : v) {
            if(it.m < 0 || it.n < 0 || it.k < 0) throw xh::Error(XH_EINVAL, "gemm plan: negative extent");
            if(it.a_off < 0 || it.b_off < 0 || it.c_off < 0) throw xh::Error(XH_EINVAL, "gemm plan: negative offset");
            if(it.m == 0 || it.n == 0 || is_skinny(it)) continue;
            const int tm = is_big(it) ? BM2 : BM, tn = is_big(it) ? BN2 : BN;
            base_tiles += int64_t((it.m + tm - 1) / tm) * ((it.n + tn - 1) / tn);
        }
        // K split: aim at 4 tiles per CU, chunks of at least 128
        const int64_t target = 4 * int64_t(std::max(ctx->num_cu, 1));
        const int64_t S = base_tiles > 0 && base_tiles < target ? (target + base_tiles - 1) / base_tiles : 1;
        std::vector<ItemExt> ext(v.size(), ItemExt{1, 0, 0});
        std::vector<RedItem> red;
        std::vector<Tile> t_big, t_rest, t_sk;
        std::vector<double> w_big, w_rest;
        for(int id : order) {
            const xh_gemm_item& it = v[id];
            if(it.m == 0 || it.n == 0) continue;
            plan->flops += 2.0 * it.m * double(it.n) * it.k;
            ItemExt& e = ext[id];
            e.kc = std::max(it.k, 1);
            if(is_skinny(it)) {  // never split: K <= SK_KMAX; blocks are made below
                plan->any_tiles = true;
                continue;
            }
            if(S > 1 && it.k >= 256) {
                int kc = static_cast<int>((it.k + S - 1) / S);
                kc = std::max(128, (kc + 15) / 16 * 16);
                const int split = (it.k + kc - 1) / kc;
                if(split > 1) {
                    e.split = split;
                    e.kc = kc;
                    e.ws_off = plan->ws_doubles;
                    const int64_t mn = int64_t(it.m) * it.n;
                    red.push_back(RedItem{it.c_off, e.ws_off, mn, split, 0});
                    plan->ws_doubles += mn * split;
                    plan->max_red_mn = std::max(plan->max_red_mn, mn);
                }
            }
            const bool big = is_big(it);
            for(int ks = 0; ks < e.split; ++ks) {
                const double kw = std::min(it.k - ks * e.kc, e.kc) + 8.0;  // + per-tile overhead
                plan->any_tiles = true;
                if(big) {
                    for(int n0 = 0; n0 < it.n; n0 += BN2)
                        for(int m0 = 0; m0 < it.m; m0 += BM2) {
                            t_big.push_back(Tile{id, m0, n0, ks});
                            w_big.push_back(4.0 * kw);
                        }
                } else {
                    for(int n0 = 0; n0 < it.n; n0 += BN)
                        for(int m0 = 0; m0 < it.m; m0 += BM) {
                            t_rest.push_back(Tile{id, m0, n0, ks});
                            w_rest.push_back(kw);
                        }
                }
            }
        }
        // skinny items: blocks of SK_THREADS * rows-per-thread long-side indices
        for(const auto& it : v)
            if(it.m > 0 && it.n > 0 && is_skinny(it))
                while(plan->sk_ms < std::min(it.m, it.n)) plan->sk_ms *= 2;
        for(int id : order) {
            const xh_gemm_item& it = v[id];
            if(it.m == 0 || it.n == 0 || !is_skinny(it)) continue;
            const int step = SK_THREADS * sk_rows(plan->sk_ms);
            for(int l0 = 0; l0 < std::max(it.m, it.n); l0 += step) t_sk.push_back(Tile{id, l0, 0, 0});
        }
        if(t_big.size() + t_rest.size() + t_sk.size() > 0x7fffffff) throw xh::Error(XH_EINVAL, "gemm plan: too many tiles");
        plan->skinny.ntiles = static_cast<int32_t>(t_sk.size());
        cut_parts(w_big, plan->big128.part);
        cut_parts(w_rest, plan->rest64.part);
        plan->big128.ntiles = static_cast<int32_t>(t_big.size());
        plan->rest64.ntiles = static_cast<int32_t>(t_rest.size());
        plan->nred = static_cast<int32_t>(red.size());
        plan->ta = trans_a ? 1 : 0;
        plan->tb = trans_b ? 1 : 0;
        {
            auto lg = [](double x) {
                uint64_t l = 0;
                while(x > 1.0 && l < 63) {
                    x *= 0.5;
                    ++l;
                }
                return l;
            };
            double mm = 0, mn = 0, mk = 0;
            for(const auto& it : v) {
                mm = std::max(mm, double(it.m));
                mn = std::max(mn, double(it.n));
                mk = std::max(mk, double(it.k));
            }
            plan->sig = (uint64_t(plan->ta) << 62) | (uint64_t(plan->tb) << 61) | (lg(double(v.size())) << 48) | (lg(mm) << 40) | (lg(mn) << 32) |
                        (lg(mk) << 24) | (lg(plan->flops) << 16) | (plan->ws_doubles > 0 ? 1u : 0u);
        }
        xh::PackedUpload& up = plan->up;
        const size_t bytes = xh::PackedUpload::packed_bytes({v.size() * sizeof(xh_gemm_item), ext.size() * sizeof(ItemExt), t_big.size() * sizeof(Tile),
                                                             t_rest.size() * sizeof(Tile), red.size() * sizeof(RedItem), t_sk.size() * sizeof(Tile)});
        if(bytes >= (size_t(1) << 18)) up.begin_pinned(ctx, bytes);  // large plans: straight into pinned staging
        auto put = [&up](const auto& vec) {
            using T = typename std::decay_t<decltype(vec)>::value_type;
            const size_t o = up.reserve<T>(static_cast<int64_t>(vec.size()));
            if(!vec.empty()) std::memcpy(up.host_ptr<T>(o), vec.data(), vec.size() * sizeof(T));
            return o;
        };
        const size_t oi = put(v), oe = put(ext), ob = put(t_big), orr_t = put(t_rest), orr = put(red), osk = put(t_sk);
        plan->up.upload(ctx);
        plan->d_items = plan->up.ptr<xh_gemm_item>(oi);
        plan->d_ext = plan->up.ptr<ItemExt>(oe);
        plan->h_items = std::move(v);
        plan->h_ext = std::move(ext);
        plan->h_order = std::move(order);
        plan->big128.d_tiles = plan->up.ptr<Tile>(ob);
        plan->rest64.d_tiles = plan->up.ptr<Tile>(orr_t);
        plan->skinny.d_tiles = plan->up.ptr<Tile>(osk);
        plan->d_red = red.empty() ? nullptr : plan->up.ptr<RedItem>(orr);
    });
    if(rc != XH_OK) {
        xh_gemm_plan_destroy(ctx, plan);
        return rc;
    }
    *out = plan;
    return XH_OK;
}

namespace {
void gemm_plan_release(xh_ctx* ctx, xh::PlanLife* p)
{
    auto* plan = static_cast<xh_gemm_plan*>(p);
    plan->up.release(ctx);
    plan->up_all.release(ctx);
    delete plan;
}
}  // namespace

extern "C" int xh_gemm_plan_destroy(xh_ctx* ctx, xh_gemm_plan* plan)
{
    if(!plan) return XH_OK;
    return guard(ctx, [&] {
        if(plan->pins > 0) {
            plan->dead = true;  // a live graph reads its tables: freed by the last xh_graph_destroy
            return;
        }
        gemm_plan_release(ctx, plan);
    });
}

extern "C" int xh_gemm_plan_flops(const xh_gemm_plan* plan, double* flops)
{
    if(!plan || !flops) return XH_EINVAL;
    *flops = plan->flops;
    return XH_OK;
}

namespace {
// the every-item 64 x 64 list, needed only by the FP32 / BF16x3 modes and the 64-only FP64 tiling
void build_all64(xh_ctx* ctx, xh_gemm_plan* plan)
{
    std::vector<Tile> t;
    std::vector<double> w;
    for(int id : plan->h_order) {
        const xh_gemm_item& it = plan->h_items[id];
        if(it.m == 0 || it.n == 0) continue;
        const ItemExt& e = plan->h_ext[id];
        for(int ks = 0; ks < e.split; ++ks) {
            const double kw = std::min(it.k - ks * e.kc, e.kc) + 8.0;
            for(int n0 = 0; n0 < it.n; n0 += BN)
                for(int m0 = 0; m0 < it.m; m0 += BM) {
                    t.push_back(Tile{id, m0, n0, ks});
                    w.push_back(kw);
                }
        }
    }
    if(t.size() > 0x7fffffff) throw xh::Error(XH_EINVAL, "gemm plan: too many tiles");
    cut_parts(w, plan->all64.part);
    plan->all64.ntiles = static_cast<int32_t>(t.size());
    const size_t o = plan->up_all.add(t);
    plan->up_all.upload(ctx);
    plan->all64.d_tiles = plan->up_all.ptr<Tile>(o);
    plan->all64_built = true;
}

// XH_OZAKI takes the int8 path for items whose output sides and K are at least OZ_MIN (the matrix
// cores' tiles are then mostly full); skinny items keep the FP64 skinny kernel and the others the
// 64 x 64 FP64 tiles, in the same call.  Both are FP64-accurate.
constexpr int OZ_MIN = 64;
bool oz_item(const xh_gemm_item& it) { return it.m >= OZ_MIN && it.n >= OZ_MIN && it.k >= OZ_MIN; }

void build_oz(xh_ctx* ctx, xh_gemm_plan* plan)
{
    std::vector<OzItem> v(plan->h_items.size(), OzItem{});
    std::vector<OzJob> jr, jk;
    int64_t bytes = 0, exps = 0;
    for(size_t id = 0; id < plan->h_items.size(); ++id) {
        const xh_gemm_item& it = plan->h_items[id];
        if(!oz_item(it)) continue;
        OzItem& o = v[id];
        o.kp = (it.k + OZ_KB - 1) / OZ_KB * OZ_KB;
        o.sa_off = bytes;
        bytes += int64_t(OZ_NS) * it.m * o.kp;
        o.sb_off = bytes;
        bytes += int64_t(OZ_NS) * it.n * o.kp;
        o.ea_off = exps;
        exps += it.m;
        o.eb_off = exps;
        exps += it.n;
        for(int op = 0; op < 2; ++op) {
            const int rows = op == 0 ? it.m : it.n;
            const bool row_contig = op == 0 ? !plan->ta : plan->tb;
            if(row_contig) {
                for(int r0 = 0; r0 < rows; r0 += 256) jr.push_back(OzJob{int32_t(id), op, r0, 1});
            } else {
                int gsz = 4;
                while(gsz < 64 && gsz < it.k) gsz *= 2;
                for(int r0 = 0; r0 < rows; r0 += 256 / gsz) jk.push_back(OzJob{int32_t(id), op, r0, gsz});
            }
        }
    }
    std::vector<Tile> t_oz, t_rest;
    std::vector<double> w_oz, w_rest;
    for(int id : plan->h_order) {
        const xh_gemm_item& it = plan->h_items[id];
        if(it.m == 0 || it.n == 0 || is_skinny(it)) continue;
        const ItemExt& e = plan->h_ext[id];
        const bool oz = oz_item(it);
        for(int ks = 0; ks < e.split; ++ks) {
            const double kw = std::min(it.k - ks * e.kc, e.kc) + 8.0;
            for(int n0 = 0; n0 < it.n; n0 += BN)
                for(int m0 = 0; m0 < it.m; m0 += BM) {
                    (oz ? t_oz : t_rest).push_back(Tile{id, m0, n0, ks});
                    (oz ? w_oz : w_rest).push_back(kw);
                }
        }
    }
    if(t_oz.size() > 0x7fffffff || t_rest.size() > 0x7fffffff || jr.size() > 0x7fffffff || jk.size() > 0x7fffffff)
        throw xh::Error(XH_EINVAL, "gemm plan: too many tiles or slicing jobs");
    cut_parts(w_oz, plan->oz_tiles.part);
    cut_parts(w_rest, plan->oz_rest.part);
    plan->oz_tiles.ntiles = static_cast<int32_t>(t_oz.size());
    plan->oz_rest.ntiles = static_cast<int32_t>(t_rest.size());
    const size_t ov = plan->up_oz.add(v), oto = plan->up_oz.add(t_oz), otr = plan->up_oz.add(t_rest), ojr = plan->up_oz.add(jr), ojk = plan->up_oz.add(jk);
    plan->up_oz.upload(ctx);
    plan->d_oz = plan->up_oz.ptr<OzItem>(ov);
    plan->oz_tiles.d_tiles = plan->up_oz.ptr<Tile>(oto);
    plan->oz_rest.d_tiles = plan->up_oz.ptr<Tile>(otr);
    plan->d_jobs_rows = plan->up_oz.ptr<OzJob>(ojr);
    plan->d_jobs_k = plan->up_oz.ptr<OzJob>(ojk);
    plan->n_jobs_rows = static_cast<int32_t>(jr.size());
    plan->n_jobs_k = static_cast<int32_t>(jk.size());
    plan->oz_slice_bytes = bytes;
    plan->oz_exps = exps;
    plan->oz_built = true;
}
}  // namespace

// ---------------------------------------------------------------- allocation-free variant
// Split-K partials of xh_gemm_noalloc live in this fixed array (one copy per device, held until
// the process ends) instead of a per-call stream-ordered allocation.  Host code must never pass
// `scratch` itself to a kernel: its device address is taken once per device with
// hipGetSymbolAddress (gemm_scratch_init, called by xh_create) and kept in the context.
static __device__ __align__(256) double scratch[XH_GEMM_SCRATCH_BYTES / sizeof(double)];

namespace xh {
void gemm_scratch_init(xh_ctx* ctx)
{
    static std::mutex mu;
    static std::unordered_map<int, double*> addr;  // device -> address of its copy of `scratch`
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(ctx->device);
    if(it == addr.end()) {
        void* p = nullptr;  // the caller has made ctx->device current
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(gemm scratch)");
        if(!p) throw Error(XH_EHIP, "hipGetSymbolAddress(gemm scratch) returned a null address");
        it = addr.emplace(ctx->device, static_cast<double*>(p)).first;
    }
    ctx->gemm_scratch = it->second;
}
}  // namespace xh

namespace {
// the log record of a timed launch: tag, ta, tb, then (m, n, k) per item
std::vector<int32_t> launch_items(const xh_ctx* ctx, const xh_gemm_plan* plan)
{
    std::vector<int32_t> v{ctx->gemm_tag, plan->ta, plan->tb};
    v.reserve(2 + 3 * plan->h_items.size());
    for(const auto& it : plan->h_items) {
        v.push_back(it.m);
        v.push_back(it.n);
        v.push_back(it.k);
    }
    return v;
}

// One grouped GEMM call.  ws: split-K partials (plan->ws_doubles, or null when the plan has no
// split items).  With noalloc the call allocates nothing: no deferred tile list is built, and no
// tuning or timing events are used.
void gemm_run(xh_ctx* ctx, const xh_gemm_plan* cplan, const double* A, const double* B, double* C, double alpha, double beta, int precision, double* ws,
              bool noalloc)
{
    {
        if(!cplan) throw xh::Error(XH_EINVAL, "null gemm plan");
        xh_gemm_plan* plan = const_cast<xh_gemm_plan*>(cplan);  // the deferred tile list is filled in on first use
        if(precision != XH_FP64 && precision != XH_FP32 && precision != XH_BF16X3 && precision != XH_OZAKI) throw xh::Error(XH_EINVAL, "unknown precision mode");
        if(precision == XH_OZAKI && noalloc) throw xh::Error(XH_EINVAL, "xh_gemm_noalloc: XH_OZAKI needs a slice workspace; use xh_gemm");
        if(!plan->any_tiles) return;
        const bool need_all = (precision != XH_FP64 && precision != XH_OZAKI) || ctx->gemm_tiles != 0;
        if(need_all && !plan->all64_built) {
            if(noalloc)
                throw xh::Error(XH_EINVAL, "xh_gemm_noalloc: the plan's 64 x 64 tile list (FP32/BF16x3 modes, 64 x 64 FP64 tiling) does not exist yet; "
                                           "one xh_gemm call in this mode builds it");
            build_all64(ctx, plan);
        }
        bool own_ws = false;
        if(plan->ws_doubles > 0 && !ws) {
            xh::refuse_under_capture(ctx, "a split-K GEMM plan (it allocates its partials; xh_gemm_noalloc does not)");
            check(xh::dmalloc(ctx, reinterpret_cast<void**>(&ws), size_t(plan->ws_doubles) * sizeof(double)), "hipMallocAsync(split K)");
            own_ws = true;
        }
        if(plan->dead) throw xh::Error(XH_EINVAL, "gemm plan used after xh_gemm_plan_destroy");
        xh::pin_for_capture(ctx, plan, gemm_plan_release);
        GemmArgs g{plan->d_items, plan->d_ext, nullptr, {}, plan->ta, plan->tb, A, B, C, ws, alpha, beta};
        // a multiple of NPART workgroups, at most per_cu per CU
        auto launch = [&](const TileList& L, void (*kernel)(GemmArgs), int per_cu, int threads, const char* name) {
            if(L.ntiles == 0) return;
            g.tiles = L.d_tiles;
            for(int x = 0; x <= NPART; ++x) g.part[x] = L.part[x];
            int grid = std::min<int>(L.ntiles, ctx->num_cu * per_cu);
            grid = (grid + NPART - 1) / NPART * NPART;
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, ctx->stream, g);
            xh::launch_check(name);
        };
        // tiles: 0 = FP64 128x128 tiles for the large items and 64x64 for the rest, 1 = 64x64 only
        auto run = [&](int tiles) {
            if(precision == XH_FP64 && tiles == 0) {
                if(plan->skinny.ntiles > 0) {
                    auto* kern = plan->sk_ms <= 8 ? k_gemm_skinny<8, sk_rows(8)> : plan->sk_ms <= 16 ? k_gemm_skinny<16, sk_rows(16)> : k_gemm_skinny<32, sk_rows(32)>;
                    const int grid = std::min<int>(plan->skinny.ntiles, ctx->num_cu * 8);
                    hipLaunchKernelGGL(kern, dim3(grid), dim3(SK_THREADS), 0, ctx->stream, g, plan->skinny.d_tiles, plan->skinny.ntiles);
                    xh::launch_check("k_gemm_skinny");
                }
                launch(plan->big128, k_gemm_f64_128, 2, G8_THREADS, "k_gemm_f64_128");
                launch(plan->rest64, k_gemm_f64p<64, 64, 16, 2, 2>, 8, G_THREADS, "k_gemm_f64");
            } else if(precision == XH_FP64) {
                launch(plan->all64, k_gemm_f64p<64, 64, 16, 2, 2>, 8, G_THREADS, "k_gemm_f64");
            } else if(precision == XH_OZAKI) {
                if(!plan->oz_built) build_oz(ctx, plan);
                if(plan->skinny.ntiles > 0) {
                    auto* kern = plan->sk_ms <= 8 ? k_gemm_skinny<8, sk_rows(8)> : plan->sk_ms <= 16 ? k_gemm_skinny<16, sk_rows(16)> : k_gemm_skinny<32, sk_rows(32)>;
                    const int grid = std::min<int>(plan->skinny.ntiles, ctx->num_cu * 8);
                    hipLaunchKernelGGL(kern, dim3(grid), dim3(SK_THREADS), 0, ctx->stream, g, plan->skinny.d_tiles, plan->skinny.ntiles);
                    xh::launch_check("k_gemm_skinny");
                }
                launch(plan->oz_rest, k_gemm_f64p<64, 64, 16, 2, 2>, 8, G_THREADS, "k_gemm_f64");
                if(plan->oz_tiles.ntiles > 0) {
                    int8_t* sl = nullptr;
                    const size_t ex_off = (size_t(plan->oz_slice_bytes) + 255) / 256 * 256;
                    const size_t bytes = ex_off + size_t(plan->oz_exps) * sizeof(int32_t);
                    xh::refuse_under_capture(ctx, "an XH_OZAKI GEMM (it allocates its slices)");
                    check(xh::dmalloc(ctx, reinterpret_cast<void**>(&sl), bytes), "hipMallocAsync(Ozaki slices)");
                    OzArgs oa{plan->d_oz, sl, reinterpret_cast<int32_t*>(sl + ex_off)};
                    if(plan->n_jobs_rows > 0) {
                        hipLaunchKernelGGL(k_oz_split_rows, dim3(static_cast<unsigned>(plan->n_jobs_rows)), dim3(256), 0, ctx->stream, g, oa, plan->d_jobs_rows);
                        xh::launch_check("k_oz_split_rows");
                    }
                    if(plan->n_jobs_k > 0) {
                        hipLaunchKernelGGL(k_oz_split_k, dim3(static_cast<unsigned>(plan->n_jobs_k)), dim3(256), 0, ctx->stream, g, oa, plan->d_jobs_k);
                        xh::launch_check("k_oz_split_k");
                    }
                    const TileList& L = plan->oz_tiles;
                    g.tiles = L.d_tiles;
                    for(int x = 0; x <= NPART; ++x) g.part[x] = L.part[x];
                    int grid = std::min<int>(L.ntiles, ctx->num_cu);
                    grid = (grid + NPART - 1) / NPART * NPART;
                    hipLaunchKernelGGL(k_gemm_oz, dim3(grid), dim3(OZ_THREADS), 0, ctx->stream, g, oa);
                    xh::launch_check("k_gemm_oz");
                    check(xh::dfree(ctx, sl), "hipFreeAsync(Ozaki slices)");
                }
            } else {
                launch(plan->all64, precision == XH_FP32 ? k_gemm_f32 : k_gemm_bf16x3, 8, G_THREADS, "k_gemm");
            }
            if(plan->nred > 0) {
                const unsigned gx = static_cast<unsigned>(std::min<int64_t>((plan->max_red_mn + 1023) / 1024, 256));
                hipLaunchKernelGGL(k_splitk_reduce, dim3(std::max(gx, 1u), static_cast<unsigned>(plan->nred)), dim3(256), 0, ctx->stream, plan->d_red, ws, C,
                                   alpha, beta);
                xh::launch_check("k_splitk_reduce");
            }
        };
        int tiles = ctx->gemm_tiles;
        const bool tunable = !noalloc && ctx->gemm_autotune && precision == XH_FP64 && plan->big128.ntiles > 0 && !xh_graph_capturing(ctx);
        const bool timed = !noalloc && ctx->gemm_timing;
        if(timed) xh::gemm_timer_begin(ctx);
        if(tunable) {
            auto hit = ctx->gemm_tune.find(plan->sig);
            if(hit != ctx->gemm_tune.end()) {
                tiles = hit->second;
                ++ctx->gemm_tune_hits;
            } else if(beta == 0.0 && C != A && C != B) {
                // time both tilings on this call's operands (beta = 0: every run rewrites C) and
                // run the faster one last, so that C holds exactly its result
                if(!plan->all64_built) build_all64(ctx, plan);
                hipEvent_t ev[3];
                for(auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
                check(hipEventRecord(ev[0], ctx->stream), "hipEventRecord");
                run(0);
                check(hipEventRecord(ev[1], ctx->stream), "hipEventRecord");
                run(1);
                check(hipEventRecord(ev[2], ctx->stream), "hipEventRecord");
                check(hipEventSynchronize(ev[2]), "hipEventSynchronize");
                float t0 = 0.f, t1 = 0.f;
                check(hipEventElapsedTime(&t0, ev[0], ev[1]), "hipEventElapsedTime");
                check(hipEventElapsedTime(&t1, ev[1], ev[2]), "hipEventElapsedTime");
                for(auto& e : ev) hipEventDestroy(e);
                tiles = t0 < t1 ? 0 : 1;
                ctx->gemm_tune[plan->sig] = tiles;
                ++ctx->gemm_tune_runs;
                if(tiles == 0) run(0);
                if(timed) xh::gemm_timer_end(ctx, plan->flops, launch_items(ctx, plan));
                if(own_ws) xh::dev_free(ctx, ws);
                return;
            }
        }
        if(tiles != 0 && !plan->all64_built) build_all64(ctx, plan);  // autotuned choice (never with noalloc)
        run(tiles);
        if(timed) xh::gemm_timer_end(ctx, plan->flops, launch_items(ctx, plan));
        if(own_ws) xh::dev_free(ctx, ws);
    }
}
}  // namespace

extern "C" int xh_gemm(xh_ctx* ctx, const xh_gemm_plan* plan, const double* A, const double* B, double* C, double alpha, double beta, int precision)
{
    return guard(ctx, [&] { gemm_run(ctx, plan, A, B, C, alpha, beta, precision, nullptr, false); });
}

extern "C" int xh_gemm_noalloc(xh_ctx* ctx, const xh_gemm_plan* plan, const double* A, const double* B, double* C, double alpha, double beta,
                               int precision)
{
    return guard(ctx, [&] {
        if(!plan) throw xh::Error(XH_EINVAL, "null gemm plan");
        const size_t need = size_t(plan->ws_doubles) * sizeof(double);
        if(need > XH_GEMM_SCRATCH_BYTES)
            throw xh::Error(XH_EINVAL, "xh_gemm_noalloc: the plan's split-K partials need " + std::to_string(need) + " bytes, more than the " +
                                           std::to_string(size_t(XH_GEMM_SCRATCH_BYTES)) + "-byte scratch array");
        if(need > 0 && !ctx->gemm_scratch) throw xh::Error(XH_EHIP, "xh_gemm_noalloc: the scratch array's device address is not set");
        gemm_run(ctx, plan, A, B, C, alpha, beta, precision, need > 0 ? ctx->gemm_scratch : nullptr, true);
    });
}

extern "C" int xh_gemm_plan_workspace(const xh_gemm_plan* plan, int64_t* bytes)
{
    if(!plan || !bytes) return XH_EINVAL;
    *bytes = plan->ws_doubles * int64_t(sizeof(double));
    return XH_OK;
}

extern "C" int xh_set_gemm_autotune(xh_ctx* ctx, int enable)
{
    return guard(ctx, [&] { ctx->gemm_autotune = enable != 0; });
}

extern "C" int xh_gemm_autotune_stats(xh_ctx* ctx, int64_t* entries, int64_t* tuned, int64_t* hits, int64_t* chose_64)
{
    return guard(ctx, [&] {
        int64_t c64 = 0;
        for(const auto& kv : ctx->gemm_tune) c64 += kv.second == 1;
        if(entries) *entries = static_cast<int64_t>(ctx->gemm_tune.size());
        if(tuned) *tuned = ctx->gemm_tune_runs;
        if(hits) *hits = ctx->gemm_tune_hits;
        if(chose_64) *chose_64 = c64;
    });
}

extern "C" int xh_set_gemm_tiles(xh_ctx* ctx, int mode)
{
    return guard(ctx, [&] {
        if(mode < 0 || mode > 1) throw xh::Error(XH_EINVAL, "gemm tile mode must be 0 or 1");
        ctx->gemm_tiles = mode;
    });
}

// ---------------------------------------------------------------- gather GEMM (permute fused into the product)

struct xh_gather_plan : xh::PlanLife {
    xh::PackedUpload up;
    GatherDev dev{};
    const xh_gemm_plan* gemm = nullptr;
    int64_t src_elems = 0;
    // host copies for the checks of a scatter plan: per item its domain tree starts (offset in
    // h_dstarts, ndom + 1 entries) and per tree the domain legs' extents (offset in h_dext)
    std::vector<int64_t> h_dstarts_off, h_dext_off;
    std::vector<int32_t> h_ndom, h_dstarts;
    std::vector<uint32_t> h_dext;
};

struct xh_scatter_plan : xh::PlanLife {
    xh::PackedUpload up;
    ScatterDev dev{};
    const xh_gather_plan* gather = nullptr;
    int64_t dst_elems = 0;
};

namespace {
bool gather_eligible(const xh_gemm_plan* plan)
{
    if(plan->ta != 0 || plan->skinny.ntiles == 0) return false;
    for(const auto& it : plan->h_items)
        if(it.m > 0 && it.n > 0 && !(is_skinny(it) && it.m > it.n)) return false;
    return true;
}

void gather_plan_release(xh_ctx* ctx, xh::PlanLife* p)
{
    auto* plan = static_cast<xh_gather_plan*>(p);
    plan->up.release(ctx);
    delete plan;
}
}  // namespace

extern "C" int xh_gemm_gather_eligible(const xh_gemm_plan* plan, int* ok)
{
    if(!plan || !ok) return XH_EINVAL;
    *ok = gather_eligible(plan) ? 1 : 0;
    return XH_OK;
}

extern "C" int xh_gather_plan_create(xh_ctx* ctx, const xh_gemm_plan* gemm, const xh_gather_desc* d, xh_gather_plan** out)
{
    *out = nullptr;
    auto* plan = new xh_gather_plan();
    const int rc = guard(ctx, [&] {
        xh::refuse_under_capture(ctx, "plan creation");
        if(!gemm || !d) throw xh::Error(XH_EINVAL, "gather plan: null argument");
        if(!gather_eligible(gemm)) throw xh::Error(XH_EINVAL, "gather plan: the GEMM plan has an item that is not a skinny item with A the long operand");
        const xh_permute_desc& pd = d->perm;
        const int NL = pd.nlegs, NR = pd.new_rank;
        if(NL < 1 || NL > XH_MAXLEGS || NR < 0 || NR > NL || NR > GA_MAXR) throw xh::Error(XH_EINVAL, "gather plan: invalid leg counts");
        const int64_t NT = pd.ntasks;
        if(NT < 0 || NT > 0x7fffffff || pd.nsources < 0) throw xh::Error(XH_EINVAL, "gather plan: invalid task or source count");
        if(d->nitems != static_cast<int64_t>(gemm->h_items.size())) throw xh::Error(XH_EINVAL, "gather plan: one gather item per GEMM item");
        if(d->nstarts < 0 || d->src_elems < 0) throw xh::Error(XH_EINVAL, "gather plan: invalid sizes");
        // every task: one source, coefficient +-1, extents matching its subblock, every source
        // element inside src
        xp::parallel_ranges(size_t(NT), 4096, [&](size_t t0, size_t t1) {
            for(size_t t = t0; t < t1; ++t) {
                const xh_permute_task& tk = pd.tasks[t];
                if(tk.nsrc != 1 || tk.src_begin < 0 || tk.src_begin >= pd.nsources) throw xh::Error(XH_EINVAL, "gather plan: every task needs exactly one source");
                const xh_permute_source& so = pd.sources[tk.src_begin];
                if(so.coeff != 1.0 && so.coeff != -1.0) throw xh::Error(XH_EINVAL, "gather plan: source coefficients must be +1 or -1");
                int64_t rprod = 1, cprod = 1, rmax = 0, cmax = 0;
                for(int k = 0; k < NL; ++k) {
                    if(tk.ext[k] == 0 || tk.sstride[k] < 0 || tk.sstride[k] >= (int64_t(1) << 31)) throw xh::Error(XH_EINVAL, "gather plan: invalid leg");
                    (k < NR ? rprod : cprod) *= tk.ext[k];
                    (tk.is_col[k] ? cmax : rmax) += (int64_t(tk.ext[k]) - 1) * tk.sstride[k];
                }
                if(rprod != tk.rows || cprod != tk.cols || tk.rows >= (int64_t(1) << 31) || tk.cols >= (int64_t(1) << 31))
                    throw xh::Error(XH_EINVAL, "gather plan: leg extents do not match the subblock shape");
                if(so.src_off < 0 || rmax >= so.src_ld || rmax >= (int64_t(1) << 31) || cmax >= (int64_t(1) << 31) || so.src_off + rmax + cmax * so.src_ld >= d->src_elems)
                    throw xh::Error(XH_EINVAL, "gather plan: source subblock out of bounds");
            }
        });
        // every item: its trees tile the A block, and its tasks have the trees' shapes
        xp::parallel_ranges(size_t(d->nitems), 1, [&](size_t j0, size_t j1) {
            for(size_t j = j0; j < j1; ++j) {
                const xh_gemm_item& it = gemm->h_items[j];
                const xh_gather_item& gi = d->items[j];
                if(it.m == 0 || it.n == 0) continue;
                if(gi.ndom < 1 || gi.ncod < 1 || gi.task_base < 0 || int64_t(gi.task_base) + int64_t(gi.ndom) * gi.ncod > NT || gi.starts_off < 0 ||
                   gi.starts_off + gi.ndom + gi.ncod + 2 > d->nstarts)
                    throw xh::Error(XH_EINVAL, "gather plan: item tables out of range");
                const int32_t* ds = d->starts + gi.starts_off;
                const int32_t* cs = ds + gi.ndom + 1;
                if(ds[0] != 0 || ds[gi.ndom] != it.m || cs[0] != 0 || cs[gi.ncod] != it.k) throw xh::Error(XH_EINVAL, "gather plan: tree starts do not cover the A block");
                if(gi.ncod > 0xffff) throw xh::Error(XH_EINVAL, "gather plan: too many codomain trees");
                // the kernel decodes a row with the domain-leg extents of one task of its tree
                for(int a = 0; a < gi.ndom; ++a) {
                    const xh_permute_task& t0 = pd.tasks[gi.task_base + int64_t(a) * gi.ncod];
                    for(int b = 0; b < gi.ncod; ++b) {
                        const xh_permute_task& tk = pd.tasks[gi.task_base + int64_t(a) * gi.ncod + b];
                        if(ds[a + 1] - ds[a] != tk.rows || cs[b + 1] - cs[b] != tk.cols) throw xh::Error(XH_EINVAL, "gather plan: task shape does not match its trees");
                        for(int k = 0; k < NR; ++k)
                            if(tk.ext[k] != t0.ext[k]) throw xh::Error(XH_EINVAL, "gather plan: domain-leg extents differ within a domain tree");
                    }
                }
            }
        });
        plan->h_ndom.assign(size_t(d->nitems), 0);
        plan->h_dstarts_off.assign(size_t(d->nitems), 0);
        plan->h_dext_off.assign(size_t(d->nitems), 0);
        for(int64_t j = 0; j < d->nitems; ++j) {
            const xh_gemm_item& it = gemm->h_items[j];
            const xh_gather_item& g = d->items[j];
            if(it.m == 0 || it.n == 0) continue;
            plan->h_ndom[j] = g.ndom;
            plan->h_dstarts_off[j] = int64_t(plan->h_dstarts.size());
            plan->h_dext_off[j] = int64_t(plan->h_dext.size());
            plan->h_dstarts.insert(plan->h_dstarts.end(), d->starts + g.starts_off, d->starts + g.starts_off + g.ndom + 1);
            for(int a = 0; a < g.ndom; ++a)
                for(int k = 0; k < NR; ++k) plan->h_dext.push_back(pd.tasks[g.task_base + int64_t(a) * g.ncod].ext[k]);
        }
        // blocks of at most SK_THREADS rows meeting at most GA_SEG domain trees, in the GEMM plan's
        // item order; per item the codomain tree and in-tree column of every column
        std::vector<GatherBlock> blocks;
        std::vector<GatherItem> items(size_t(d->nitems), GatherItem{0, 0, 0});
        std::vector<int32_t> cols;
        for(int64_t j = 0; j < d->nitems; ++j) {
            const xh_gemm_item& it = gemm->h_items[j];
            const xh_gather_item& g = d->items[j];
            if(it.m == 0 || it.n == 0) continue;
            const int32_t* ds = d->starts + g.starts_off;
            const int32_t* cs = ds + g.ndom + 1;
            items[j] = GatherItem{g.task_base, g.ncod, int64_t(cols.size())};
            for(int b = 0; b < g.ncod; ++b)
                for(int c = cs[b]; c < cs[b + 1]; ++c) cols.push_back(b | ((c - cs[b]) << 16));
            int a = 0;
            for(int r = 0; r < it.m;) {
                GatherBlock bl{int32_t(j), r, 0, 0, {}, {}};
                const int end = std::min(it.m, r + SK_THREADS);
                while(r < end && bl.nseg < GA_SEG) {
                    while(ds[a + 1] <= r) ++a;
                    bl.a[bl.nseg] = a;
                    bl.start[bl.nseg] = ds[a];
                    ++bl.nseg;
                    r = std::min(end, ds[a + 1]);
                }
                bl.nrows = r - bl.row0;
                blocks.push_back(bl);
            }
        }
        if(blocks.size() > 0x7fffffff) throw xh::Error(XH_EINVAL, "gather plan: too many blocks");
        xh::PackedUpload& up = plan->up;
        const size_t bytes = xh::PackedUpload::packed_bytes({blocks.size() * sizeof(GatherBlock), items.size() * sizeof(GatherItem), cols.size() * sizeof(int32_t),
                                                             size_t(NT * NL) * sizeof(xh::LegInfo), size_t(NT) * sizeof(GatherSrc)});
        if(bytes >= (size_t(1) << 18)) up.begin_pinned(ctx, bytes);
        auto put = [&up](const auto& vec) {
            using T = typename std::decay_t<decltype(vec)>::value_type;
            const size_t o = up.reserve<T>(static_cast<int64_t>(vec.size()));
            if(!vec.empty()) std::memcpy(up.host_ptr<T>(o), vec.data(), vec.size() * sizeof(T));
            return o;
        };
        const size_t o_bl = put(blocks), o_it = put(items), o_co = put(cols);
        const size_t o_lg = up.reserve<xh::LegInfo>(NT * NL), o_src = up.reserve<GatherSrc>(NT);
        xh::LegInfo* lg = up.host_ptr<xh::LegInfo>(o_lg);
        GatherSrc* gs = up.host_ptr<GatherSrc>(o_src);
        xp::parallel_ranges(size_t(NT), 4096, [&](size_t t0, size_t t1) {
            for(size_t t = t0; t < t1; ++t) {
                const xh_permute_task& tk = pd.tasks[t];
                for(int k = 0; k < NL; ++k) lg[t * NL + k] = xh::make_leg(tk.ext[k], tk.sstride[k], tk.is_col[k] != 0);
                const xh_permute_source& so = pd.sources[tk.src_begin];
                gs[t] = GatherSrc{so.src_off, so.src_ld, so.coeff < 0.0 ? 1 : 0, 0};
            }
        });
        up.upload(ctx);
        plan->dev = GatherDev{up.ptr<GatherBlock>(o_bl), up.ptr<GatherItem>(o_it), up.ptr<int32_t>(o_co), up.ptr<xh::LegInfo>(o_lg), up.ptr<GatherSrc>(o_src),
                              NL, NR, static_cast<int32_t>(blocks.size())};
        plan->gemm = gemm;
        plan->src_elems = d->src_elems;
    });
    if(rc != XH_OK) {
        gather_plan_release(ctx, plan);
        return rc;
    }
    *out = plan;
    return XH_OK;
}

extern "C" int xh_gather_plan_destroy(xh_ctx* ctx, xh_gather_plan* plan)
{
    if(!plan) return XH_OK;
    return guard(ctx, [&] {
        if(plan->pins > 0) {
            plan->dead = true;  // a live graph reads its tables: freed by the last xh_graph_destroy
            return;
        }
        gather_plan_release(ctx, plan);
    });
}

extern "C" int xh_gemm_gather(xh_ctx* ctx, const xh_gemm_plan* plan, const xh_gather_plan* gp, const double* src, int64_t src_elems, const double* B, double* C,
                              double alpha, double beta)
{
    return guard(ctx, [&] {
        if(!plan || !gp) throw xh::Error(XH_EINVAL, "xh_gemm_gather: null plan");
        if(gp->gemm != plan) throw xh::Error(XH_EINVAL, "xh_gemm_gather: the gather plan was built for another GEMM plan");
        if(plan->dead || gp->dead) throw xh::Error(XH_EINVAL, "xh_gemm_gather: plan used after destroy");
        if(src_elems != gp->src_elems) throw xh::Error(XH_EINVAL, "xh_gemm_gather: the source size differs from the plan's");
        auto* mp = const_cast<xh_gemm_plan*>(plan);
        auto* mg = const_cast<xh_gather_plan*>(gp);
        xh::pin_for_capture(ctx, mp, gemm_plan_release);
        xh::pin_for_capture(ctx, mg, gather_plan_release);
        const bool timed = ctx->gemm_timing;
        if(timed) xh::gemm_timer_begin(ctx);
        GemmArgs g{plan->d_items, plan->d_ext, nullptr, {}, plan->ta, plan->tb, nullptr, B, C, nullptr, alpha, beta};
        auto* kern = plan->sk_ms <= 8 ? k_gemm_skinny_gather<8, false> : plan->sk_ms <= 16 ? k_gemm_skinny_gather<16, false> : k_gemm_skinny_gather<32, false>;
        const int grid = std::min<int>(gp->dev.nblocks, ctx->num_cu * 8);
        if(grid > 0) hipLaunchKernelGGL(kern, dim3(grid), dim3(SK_THREADS), 0, ctx->stream, g, gp->dev, src, ScatterDev{});
        xh::launch_check("k_gemm_skinny_gather");
        if(timed) xh::gemm_timer_end(ctx, plan->flops, launch_items(ctx, plan));
    });
}

// ---------------------------------------------------------------- the product's permute folded into the gather GEMM's stores

namespace {
void scatter_plan_release(xh_ctx* ctx, xh::PlanLife* p)
{
    auto* plan = static_cast<xh_scatter_plan*>(p);
    plan->up.release(ctx);
    delete plan;
}
}  // namespace

extern "C" int xh_scatter_plan_create(xh_ctx* ctx, const xh_gather_plan* gather, const xh_scatter_desc* d, xh_scatter_plan** out)
{
    *out = nullptr;
    auto* plan = new xh_scatter_plan();
    const int rc = guard(ctx, [&] {
        xh::refuse_under_capture(ctx, "plan creation");
        if(!gather || !d || !d->p) throw xh::Error(XH_EINVAL, "scatter plan: null argument");
        const xh_gemm_plan* gemm = gather->gemm;
        const xh_permute_desc& pd = d->perm;
        const int NP = pd.nlegs, NR2 = pd.new_rank, R = gather->dev.new_rank;
        if(NP < 1 || NP > XH_MAXLEGS || NR2 < 0 || NR2 > NP || R > NP) throw xh::Error(XH_EINVAL, "scatter plan: invalid leg counts");
        std::vector<int> pinv(size_t(NP), -1);
        for(int k = 0; k < NP; ++k) {
            if(d->p[k] < 0 || d->p[k] >= NP || pinv[size_t(d->p[k])] >= 0) throw xh::Error(XH_EINVAL, "scatter plan: p is not a permutation");
            pinv[size_t(d->p[k])] = k;
        }
        const int64_t NT = pd.ntasks;
        if(NT < 0 || NT > 0x7fffffff || pd.nsources < 0) throw xh::Error(XH_EINVAL, "scatter plan: invalid task or source count");
        if(d->nitems != static_cast<int64_t>(gemm->h_items.size())) throw xh::Error(XH_EINVAL, "scatter plan: one entry per GEMM item");
        if(d->ngrid < 0 || d->ncstarts < 0 || d->dst_elems < 0) throw xh::Error(XH_EINVAL, "scatter plan: invalid sizes");
        // per task: one source, coefficient +-1; its destination strides by product leg, and every
        // destination element inside dst
        std::vector<xh::LegInfo> legs(static_cast<size_t>(NT * NP));
        std::vector<int64_t> dst_off(static_cast<size_t>(NT));
        std::vector<int32_t> neg(static_cast<size_t>(NT));
        xp::parallel_ranges(size_t(NT), 4096, [&](size_t t0, size_t t1) {
            for(size_t t = t0; t < t1; ++t) {
                const xh_permute_task& tk = pd.tasks[t];
                if(tk.nsrc != 1 || tk.src_begin < 0 || tk.src_begin >= pd.nsources) throw xh::Error(XH_EINVAL, "scatter plan: every task needs exactly one source");
                const double c = pd.sources[tk.src_begin].coeff;
                if(c != 1.0 && c != -1.0) throw xh::Error(XH_EINVAL, "scatter plan: source coefficients must be +1 or -1");
                if(tk.dst_off < 0 || tk.dst_ld < 1) throw xh::Error(XH_EINVAL, "scatter plan: invalid destination subblock");
                int64_t rs = 1, cs = tk.dst_ld, dmax = tk.dst_off;
                for(int k = 0; k < NP; ++k) {
                    if(tk.ext[k] == 0) throw xh::Error(XH_EINVAL, "scatter plan: zero extent");
                    int64_t& st = k < NR2 ? rs : cs;
                    if(st >= (int64_t(1) << 31)) throw xh::Error(XH_EINVAL, "scatter plan: destination stride out of range");
                    const int j = d->p[k];  // output leg k is product leg j
                    legs[t * NP + size_t(j)] = xh::make_leg(tk.ext[k], st, false);
                    dmax += (int64_t(tk.ext[k]) - 1) * st;
                    st *= tk.ext[k];
                }
                if(dmax >= d->dst_elems || dmax >= (int64_t(1) << 62)) throw xh::Error(XH_EINVAL, "scatter plan: destination subblock out of bounds");
                dst_off[t] = tk.dst_off;
                neg[t] = c < 0.0 ? 1 : 0;
            }
        });
        // per item: its task grid covers the product block's subblocks (a, c), each task reads
        // exactly the subblock it is listed for, with the domain-leg extents the gather decodes
        std::vector<int32_t> seen(static_cast<size_t>(NT), 0);
        std::vector<ScatterItem> items(size_t(d->nitems), ScatterItem{0, 0, 0, 0});
        std::vector<int32_t> cols;
        for(int64_t j = 0; j < d->nitems; ++j) {
            const xh_gemm_item& it = gemm->h_items[j];
            if(it.m == 0 || it.n == 0) continue;
            const int ndom = gather->h_ndom[j], ncod = d->ncod[j];
            const int32_t* ds = gather->h_dstarts.data() + gather->h_dstarts_off[j];
            const uint32_t* dext = gather->h_dext.data() + gather->h_dext_off[j];
            if(ncod < 1 || ncod > 0xffff || d->grid_off[j] < 0 || d->grid_off[j] + int64_t(ndom) * ncod > d->ngrid || d->cstarts_off[j] < 0 ||
               d->cstarts_off[j] + ncod + 1 > d->ncstarts)
                throw xh::Error(XH_EINVAL, "scatter plan: item tables out of range");
            const int32_t* cs = d->cstarts + d->cstarts_off[j];
            if(cs[0] != 0 || cs[ncod] != it.n) throw xh::Error(XH_EINVAL, "scatter plan: column starts do not cover the product block");
            for(int a = 0; a < ndom; ++a)
                for(int c = 0; c < ncod; ++c) {
                    const int32_t t = d->grid[d->grid_off[j] + int64_t(a) * ncod + c];
                    if(t < 0 || t >= NT || seen[size_t(t)]++) throw xh::Error(XH_EINVAL, "scatter plan: every task must appear once in the grids");
                    const xh_permute_task& tk = pd.tasks[t];
                    const xh_permute_source& so = pd.sources[tk.src_begin];
                    if(so.src_off != it.c_off + ds[a] + int64_t(cs[c]) * it.m || so.src_ld != it.m)
                        throw xh::Error(XH_EINVAL, "scatter plan: a task does not read the product subblock it is listed for");
                    int64_t rp = 1, cp = 1;
                    for(int jj = 0; jj < NP; ++jj) {
                        const uint32_t e = tk.ext[pinv[size_t(jj)]];
                        if(jj < R) {
                            if(e != dext[int64_t(a) * R + jj]) throw xh::Error(XH_EINVAL, "scatter plan: domain-leg extents differ from the gather's");
                            rp *= e;
                        } else {
                            cp *= e;
                        }
                    }
                    if(rp != ds[a + 1] - ds[a] || cp != cs[c + 1] - cs[c]) throw xh::Error(XH_EINVAL, "scatter plan: task shape does not match its product subblock");
                }
            items[j] = ScatterItem{d->grid_off[j], int64_t(cols.size()), ncod, 0};
            for(int c = 0; c < ncod; ++c)
                for(int x = cs[c]; x < cs[c + 1]; ++x) cols.push_back(c | ((x - cs[c]) << 16));
        }
        for(int64_t t = 0; t < NT; ++t)
            if(!seen[size_t(t)]) throw xh::Error(XH_EINVAL, "scatter plan: a task is in no grid");
        xh::PackedUpload& up = plan->up;
        const size_t bytes = xh::PackedUpload::packed_bytes({items.size() * sizeof(ScatterItem), size_t(d->ngrid) * sizeof(int32_t), cols.size() * sizeof(int32_t),
                                                             legs.size() * sizeof(xh::LegInfo), dst_off.size() * sizeof(int64_t), neg.size() * sizeof(int32_t)});
        if(bytes >= (size_t(1) << 18)) up.begin_pinned(ctx, bytes);
        auto put = [&up](const auto* ptr, size_t n) {
            using T = std::decay_t<decltype(*ptr)>;
            const size_t o = up.reserve<T>(static_cast<int64_t>(n));
            if(n > 0) std::memcpy(up.host_ptr<T>(o), ptr, n * sizeof(T));
            return o;
        };
        const size_t o_it = put(items.data(), items.size()), o_gr = put(d->grid, size_t(d->ngrid)), o_co = put(cols.data(), cols.size()),
                     o_lg = put(legs.data(), legs.size()), o_do = put(dst_off.data(), dst_off.size()), o_ng = put(neg.data(), neg.size());
        up.upload(ctx);
        plan->dev = ScatterDev{up.ptr<ScatterItem>(o_it), up.ptr<int32_t>(o_gr), up.ptr<int32_t>(o_co), up.ptr<xh::LegInfo>(o_lg), up.ptr<int64_t>(o_do),
                               up.ptr<int32_t>(o_ng), NP};
        plan->gather = gather;
        plan->dst_elems = d->dst_elems;
    });
    if(rc != XH_OK) {
        scatter_plan_release(ctx, plan);
        return rc;
    }
    *out = plan;
    return XH_OK;
}

extern "C" int xh_scatter_plan_destroy(xh_ctx* ctx, xh_scatter_plan* plan)
{
    if(!plan) return XH_OK;
    return guard(ctx, [&] {
        if(plan->pins > 0) {
            plan->dead = true;  // a live graph reads its tables: freed by the last xh_graph_destroy
            return;
        }
        scatter_plan_release(ctx, plan);
    });
}

extern "C" int xh_gemm_gather_scatter(xh_ctx* ctx, const xh_gemm_plan* plan, const xh_gather_plan* gp, const xh_scatter_plan* sp, const double* src,
                                      int64_t src_elems, const double* B, double* dst, int64_t dst_elems, double alpha)
{
    return guard(ctx, [&] {
        if(!plan || !gp || !sp) throw xh::Error(XH_EINVAL, "xh_gemm_gather_scatter: null plan");
        if(gp->gemm != plan || sp->gather != gp) throw xh::Error(XH_EINVAL, "xh_gemm_gather_scatter: the plans were built for one another's others");
        if(plan->dead || gp->dead || sp->dead) throw xh::Error(XH_EINVAL, "xh_gemm_gather_scatter: plan used after destroy");
        if(src_elems != gp->src_elems || dst_elems != sp->dst_elems) throw xh::Error(XH_EINVAL, "xh_gemm_gather_scatter: buffer sizes differ from the plans'");
        xh::pin_for_capture(ctx, const_cast<xh_gemm_plan*>(plan), gemm_plan_release);
        xh::pin_for_capture(ctx, const_cast<xh_gather_plan*>(gp), gather_plan_release);
        xh::pin_for_capture(ctx, const_cast<xh_scatter_plan*>(sp), scatter_plan_release);
        const bool timed = ctx->gemm_timing;
        if(timed) xh::gemm_timer_begin(ctx);
        GemmArgs g{plan->d_items, plan->d_ext, nullptr, {}, plan->ta, plan->tb, nullptr, B, dst, nullptr, alpha, 0.0};
        auto* kern = plan->sk_ms <= 8 ? k_gemm_skinny_gather<8, true> : plan->sk_ms <= 16 ? k_gemm_skinny_gather<16, true> : k_gemm_skinny_gather<32, true>;
        const int grid = std::min<int>(gp->dev.nblocks, ctx->num_cu * 8);
        if(grid > 0) hipLaunchKernelGGL(kern, dim3(grid), dim3(SK_THREADS), 0, ctx->stream, g, gp->dev, src, sp->dev);
        xh::launch_check("k_gemm_skinny_gather");
        if(timed) xh::gemm_timer_end(ctx, plan->flops, launch_items(ctx, plan));
    });
}
