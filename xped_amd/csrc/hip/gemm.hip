// Grouped, persistent block GEMM on the CDNA4 matrix cores.
//
// The reference multiplies matched sectors one at a time with Eigen/BLAS
// (src/Core/TensorBase.cpp:262-279 -> MatrixInterface::prod, MatrixInterface_Eigen_impl.cpp:155-160).
// Here every (sector, 64x64 output tile) pair of one contraction is an entry of a single work
// list, processed by one persistent launch: workgroups stride through the list, so uneven
// sector sizes are balanced across the 256 CUs.  Tiles of one sector are spread over workgroups
// whose blockIdx are congruent mod 8, i.e. they run on the same XCD and share its L2.
//
// Precision modes (operands are always FP64 in memory, the output is FP64):
//   XH_FP64   v_mfma_f64_16x16x4_f64, FP64 accumulation
//   XH_FP32   operands rounded to FP32, v_mfma_f32_16x16x4_f32, FP32 accumulation
//   XH_BF16X3 operands rounded to FP32 and split x = hi + lo with hi, lo in BF16;
//             A*B ~ hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x16_bf16, FP32 accumulation
#include <algorithm>
#include <cstring>

#include "device.hpp"

using xh::check;
using xh::guard;

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64;
constexpr int G_THREADS = 256;

struct Tile {
    int32_t item, m0, n0, pad;
};

struct GemmArgs {
    const xh_gemm_item* items;
    const Tile* tiles;
    int32_t ntiles;
    int32_t ta, tb;
    const double* A;
    const double* B;
    double* C;
    double alpha, beta;
};

__device__ __forceinline__ double load_a(const GemmArgs& g, const xh_gemm_item& it, int i, int kk)
{
    if(i >= it.m || kk >= it.k) return 0.0;
    return g.ta ? g.A[it.a_off + int64_t(i) * it.k + kk] : g.A[it.a_off + int64_t(kk) * it.m + i];
}
__device__ __forceinline__ double load_b(const GemmArgs& g, const xh_gemm_item& it, int kk, int j)
{
    if(kk >= it.k || j >= it.n) return 0.0;
    return g.tb ? g.B[it.b_off + int64_t(kk) * it.n + j] : g.B[it.b_off + int64_t(j) * it.k + kk];
}

// Output layout of the 16x16 accumulators: for the FP32-accumulating MFMAs register r of lane l
// holds row 4*(l/16) + r; for v_mfma_f64_16x16x4_f64 it holds row (l/16) + 4*r.  Column = l % 16.
template <bool F64LAYOUT, typename Acc>
__device__ __forceinline__ void store_tile(const GemmArgs& g, const xh_gemm_item& it, int m0, int n0, int wm, int wn, int lane, Acc acc[2][2])
{
    for(int mi = 0; mi < 2; ++mi)
        for(int ni = 0; ni < 2; ++ni)
            for(int r = 0; r < 4; ++r) {
                const int rr = F64LAYOUT ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                const int row = m0 + wm * 32 + mi * 16 + rr;
                const int col = n0 + wn * 32 + ni * 16 + (lane & 15);
                if(row < it.m && col < it.n) {
                    double* p = g.C + it.c_off + int64_t(col) * it.m + row;
                    const double v = static_cast<double>(acc[mi][ni][r]);
                    *p = (g.beta == 0.0) ? g.alpha * v : g.alpha * v + g.beta * *p;
                }
            }
}

// ---------------------------------------------------------------- FP64
__global__ __launch_bounds__(G_THREADS) void k_gemm_f64(GemmArgs g)
{
    constexpr int BK = 16;
    __shared__ double sA[BK][BM + 1];
    __shared__ double sB[BK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    for(int t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        d4 acc[2][2];
        for(int a = 0; a < 2; ++a)
            for(int b = 0; b < 2; ++b) acc[a][b] = d4{0, 0, 0, 0};
        for(int k0 = 0; k0 < it.k; k0 += BK) {
            for(int r = 0; r < 4; ++r) {
                const int idx = tid + G_THREADS * r;
                const int i = idx & 63, kk = idx >> 6;
                sA[kk][i] = load_a(g, it, tl.m0 + i, k0 + kk);
                const int kb = idx & 15, j = idx >> 4;
                sB[kb][j] = load_b(g, it, k0 + kb, tl.n0 + j);
            }
            __syncthreads();
            for(int kq = 0; kq < BK; kq += 4) {
                const int kr = kq + (lane >> 4);
                double av[2], bv[2];
                for(int mi = 0; mi < 2; ++mi) av[mi] = sA[kr][wm * 32 + mi * 16 + (lane & 15)];
                for(int ni = 0; ni < 2; ++ni) bv[ni] = sB[kr][wn * 32 + ni * 16 + (lane & 15)];
                for(int mi = 0; mi < 2; ++mi)
                    for(int ni = 0; ni < 2; ++ni) acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
            }
            __syncthreads();
        }
        store_tile<true>(g, it, tl.m0, tl.n0, wm, wn, lane, acc);
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
    for(int t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        f4 acc[2][2];
        for(int a = 0; a < 2; ++a)
            for(int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
        for(int k0 = 0; k0 < it.k; k0 += BK) {
            for(int r = 0; r < 8; ++r) {
                const int idx = tid + G_THREADS * r;
                const int i = idx & 63, kk = idx >> 6;
                sA[kk][i] = static_cast<float>(load_a(g, it, tl.m0 + i, k0 + kk));
                const int kb = idx & 31, j = idx >> 5;
                sB[kb][j] = static_cast<float>(load_b(g, it, k0 + kb, tl.n0 + j));
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
        store_tile<false>(g, it, tl.m0, tl.n0, wm, wn, lane, acc);
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
    for(int t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
        const Tile tl = g.tiles[t];
        const xh_gemm_item it = g.items[tl.item];
        f4 acc[2][2];
        for(int a = 0; a < 2; ++a)
            for(int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
        for(int k0 = 0; k0 < it.k; k0 += BK) {
            for(int r = 0; r < 8; ++r) {
                const int idx = tid + G_THREADS * r;
                const int i = idx & 63, kk = idx >> 6;
                float a = static_cast<float>(load_a(g, it, tl.m0 + i, k0 + kk));
                unsigned short ah = bf16_rn(a);
                sAh[kk][i] = ah;
                sAl[kk][i] = bf16_rn(a - bf16_to_f(ah));
                const int kb = idx & 31, j = idx >> 5;
                float b = static_cast<float>(load_b(g, it, k0 + kb, tl.n0 + j));
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
        store_tile<false>(g, it, tl.m0, tl.n0, wm, wn, lane, acc);
    }
}

}  // namespace

struct xh_gemm_plan {
    xh::PackedUpload up;
    xh_gemm_item* d_items = nullptr;
    Tile* d_tiles = nullptr;
    int32_t ntiles = 0;
    int ta = 0, tb = 0;
    double flops = 0.0;
};

extern "C" int xh_gemm_plan_create(xh_ctx* ctx, const xh_gemm_item* items, int64_t nitems, int trans_a, int trans_b, xh_gemm_plan** out)
{
    *out = nullptr;
    auto* plan = new xh_gemm_plan();
    int rc = guard(ctx, [&] {
        std::vector<xh_gemm_item> v(items, items + nitems);
        // largest K*tiles first so the persistent loop ends with small work
        std::vector<int> order(v.size());
        for(size_t i = 0; i < v.size(); ++i) order[i] = static_cast<int>(i);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return double(v[a].m) * v[a].n * v[a].k > double(v[b].m) * v[b].n * v[b].k; });
        std::vector<Tile> tiles;
        for(int id : order) {
            const xh_gemm_item& it = v[id];
            if(it.m < 0 || it.n < 0 || it.k < 0) throw xh::Error(XH_EINVAL, "gemm plan: negative extent");
            if(it.m == 0 || it.n == 0) continue;
            plan->flops += 2.0 * it.m * double(it.n) * it.k;
            for(int n0 = 0; n0 < it.n; n0 += BN)
                for(int m0 = 0; m0 < it.m; m0 += BM) tiles.push_back(Tile{id, m0, n0, 0});
        }
        if(tiles.size() > 0x7fffffff) throw xh::Error(XH_EINVAL, "gemm plan: too many tiles");
        plan->ntiles = static_cast<int32_t>(tiles.size());
        plan->ta = trans_a ? 1 : 0;
        plan->tb = trans_b ? 1 : 0;
        const size_t oi = plan->up.add(v), ot = plan->up.add(tiles);
        plan->up.upload(ctx);
        plan->d_items = plan->up.ptr<xh_gemm_item>(oi);
        plan->d_tiles = plan->up.ptr<Tile>(ot);
    });
    if(rc != XH_OK) {
        xh_gemm_plan_destroy(ctx, plan);
        return rc;
    }
    *out = plan;
    return XH_OK;
}

extern "C" int xh_gemm_plan_destroy(xh_ctx* ctx, xh_gemm_plan* plan)
{
    if(!plan) return XH_OK;
    return guard(ctx, [&] {
        plan->up.release(ctx);
        delete plan;
    });
}

extern "C" int xh_gemm_plan_flops(const xh_gemm_plan* plan, double* flops)
{
    if(!plan || !flops) return XH_EINVAL;
    *flops = plan->flops;
    return XH_OK;
}

extern "C" int xh_gemm(xh_ctx* ctx, const xh_gemm_plan* plan, const double* A, const double* B, double* C, double alpha, double beta, int precision)
{
    return guard(ctx, [&] {
        if(!plan) throw xh::Error(XH_EINVAL, "null gemm plan");
        if(plan->ntiles == 0) return;
        GemmArgs g{plan->d_items, plan->d_tiles, plan->ntiles, plan->ta, plan->tb, A, B, C, alpha, beta};
        const int grid = std::min<int>(plan->ntiles, ctx->num_cu * 8);
        switch(precision) {
        case XH_FP64: hipLaunchKernelGGL(k_gemm_f64, dim3(grid), dim3(G_THREADS), 0, ctx->stream, g); break;
        case XH_FP32: hipLaunchKernelGGL(k_gemm_f32, dim3(grid), dim3(G_THREADS), 0, ctx->stream, g); break;
        case XH_BF16X3: hipLaunchKernelGGL(k_gemm_bf16x3, dim3(grid), dim3(G_THREADS), 0, ctx->stream, g); break;
        default: throw xh::Error(XH_EINVAL, "unknown precision mode");
        }
        xh::launch_check("k_gemm");
    });
}
