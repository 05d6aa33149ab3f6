// Reverse-mode building blocks: the block SVD adjoint.
//
// Per sector, with A = U S V (U m x k, S k x k diagonal, V = Vdag k x n, k = min(m, n)) and the
// incoming adjoints dU, dS, dV, the adjoint of A is (ADTensor.hpp:262-320)
//
//   dA += U (dS + (J + J^T) S + S (K + K^T)) V + (1 - U U^T) dU S^-1 V + U S^-1 dV (1 - V^T V),
//   J = F o (U^T dU),  K = F o (V dV^T),  F_ij = f(s_j^2 - s_i^2),
//   f(d) = d / (d^2 + 1e-12) for |d| < 1e-12, 1 / d otherwise  (ADTensor.hpp:279-282).
//
// With X = U^T dU and Y = V dV^T (both k x k) the projector terms fold into the k x k core,
// since U^T (dU S^-1) = X S^-1 and (S^-1 dV) V^T = S^-1 Y^T:
//
//   dA += U (M V + S^-1 dV) + (dU S^-1) V,
//   M_ij = dS_ij + F_ij (X_ij - X_ji) s_j + s_i F_ij (Y_ij - Y_ji) - X_ij / s_j - Y_ji / s_i,
//
// using F_ji = -F_ij.  The launch sequence is two grouped GEMMs (X, Y), one elementwise kernel
// (M, W = S^-1 dV, Z = dU S^-1), and three grouped GEMMs (W += M V, dA += U W, dA += Z V), so the
// m x m and n x n projectors of the reference are never formed.
#include <algorithm>
#include <vector>

#include "device.hpp"

using xh::check;
using xh::guard;

namespace {

constexpr int SB_THREADS = 256;

__device__ __forceinline__ double f_reg(double d)
{
    return fabs(d) < 1.e-12 ? d / (d * d + 1.e-12) : 1.0 / d;
}

struct SbItem {
    int64_t u_off, s_off, v_off;
    int32_t m, n, k;
};

// blockIdx.y = sector, blockIdx.z = 0: M (k x k), 1: W = S^-1 dV (k x n), 2: Z = dU S^-1 (m x k).
__global__ __launch_bounds__(SB_THREADS) void k_svd_bwd_prep(const SbItem* __restrict__ items, const double* __restrict__ S, const double* __restrict__ dS,
                                                              const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ dU,
                                                              const double* __restrict__ dV, double* __restrict__ M, double* __restrict__ W,
                                                              double* __restrict__ Z)
{
    const SbItem it = items[blockIdx.y];
    const int64_t k = it.k;
    const double* s = S + it.s_off;  // diagonal at s[i + i k]
    const int64_t stride = int64_t(gridDim.x) * SB_THREADS;
    const int64_t e0 = int64_t(blockIdx.x) * SB_THREADS + threadIdx.x;
    if(blockIdx.z == 0) {
        const double* x = X + it.s_off;
        const double* y = Y + it.s_off;
        const double* ds = dS + it.s_off;
        double* out = M + it.s_off;
        for(int64_t e = e0; e < k * k; e += stride) {
            const int64_t i = e % k, j = e / k;
            const double si = s[i + i * k], sj = s[j + j * k];
            const double F = (i == j) ? 0.0 : f_reg(sj * sj - si * si);
            const double xij = x[i + j * k], xji = x[j + i * k];
            const double yij = y[i + j * k], yji = y[j + i * k];
            out[e] = ds[e] + F * (xij - xji) * sj + si * F * (yij - yji) - xij / sj - yji / si;
        }
    } else if(blockIdx.z == 1) {
        const int64_t n = it.n;
        const double* dv = dV + it.v_off;
        double* w = W + it.v_off;
        for(int64_t e = e0; e < k * n; e += stride) {
            const int64_t i = e % k;
            w[e] = dv[e] / s[i + i * k];
        }
    } else {
        const int64_t m = it.m;
        const double* du = dU + it.u_off;
        double* z = Z + it.u_off;
        for(int64_t e = e0; e < m * k; e += stride) {
            const int64_t j = e / m;
            z[e] = du[e] / s[j + j * k];
        }
    }
}

struct GemmPlan {
    xh_ctx* ctx;
    xh_gemm_plan* p = nullptr;
    GemmPlan(xh_ctx* c, const std::vector<xh_gemm_item>& items, int ta, int tb) : ctx(c)
    {
        if(items.empty()) return;
        const int rc = xh_gemm_plan_create(ctx, items.data(), static_cast<int64_t>(items.size()), ta, tb, &p);
        if(rc != XH_OK) throw xh::Error(rc, std::string("svd adjoint plan: ") + xh_last_error(ctx));
    }
    ~GemmPlan()
    {
        if(p) xh_gemm_plan_destroy(ctx, p);
    }
    void run(const double* A, const double* B, double* C, double beta) const
    {
        if(!p) return;
        const int rc = xh_gemm(ctx, p, A, B, C, 1.0, beta, XH_FP64);
        if(rc != XH_OK) throw xh::Error(rc, std::string("svd adjoint gemm: ") + xh_last_error(ctx));
    }
};

}  // namespace

extern "C" int xh_svd_bwd(xh_ctx* ctx, const xh_svd_bwd_item* items, int64_t nitems, const double* U, const double* S, const double* V, const double* dU,
                          const double* dS, const double* dV, double* dA)
{
    return guard(ctx, [&] {
        if(nitems <= 0) return;
        std::vector<SbItem> act;
        std::vector<xh_gemm_item> gx, gy, gw, gu, gz;
        int64_t nkk = 0, nkn = 0, nmk = 0;  // extents of the S, V and U buffers touched
        int64_t max_e = 0;
        for(int64_t g = 0; g < nitems; ++g) {
            const xh_svd_bwd_item& it = items[g];
            if(it.m < 0 || it.n < 0 || it.k < 0 || it.k > std::min(it.m, it.n)) throw xh::Error(XH_EINVAL, "svd adjoint: bad block shape");
            if(it.u_off < 0 || it.s_off < 0 || it.v_off < 0 || it.a_off < 0) throw xh::Error(XH_EINVAL, "svd adjoint: negative offset");
            if(it.k == 0 || it.m == 0 || it.n == 0) continue;
            const int64_t m = it.m, n = it.n, k = it.k;
            act.push_back(SbItem{it.u_off, it.s_off, it.v_off, it.m, it.n, it.k});
            nkk = std::max(nkk, it.s_off + k * k);
            nkn = std::max(nkn, it.v_off + k * n);
            nmk = std::max(nmk, it.u_off + m * k);
            max_e = std::max({max_e, k * k, k * n, m * k});
            gx.push_back(xh_gemm_item{it.u_off, it.u_off, it.s_off, it.k, it.k, it.m});  // X = U^T dU
            gy.push_back(xh_gemm_item{it.v_off, it.v_off, it.s_off, it.k, it.k, it.n});  // Y = V dV^T
            gw.push_back(xh_gemm_item{it.s_off, it.v_off, it.v_off, it.k, it.n, it.k});  // W += M V
            gu.push_back(xh_gemm_item{it.u_off, it.v_off, it.a_off, it.m, it.n, it.k});  // dA += U W, dA += Z V
        }
        if(act.empty()) return;
        if(max_e >= (int64_t(1) << 40)) throw xh::Error(XH_EINVAL, "svd adjoint: block too large");
        double *X = nullptr, *Y = nullptr, *M = nullptr, *W = nullptr, *Z = nullptr;
        auto alloc = [&](double** p, int64_t n) { check(xh::dmalloc(ctx, reinterpret_cast<void**>(p), size_t(std::max<int64_t>(n, 1)) * sizeof(double)), "hipMallocAsync"); };
        struct Free {
            xh_ctx* c;
            double** p[5];
            ~Free()
            {
                for(double** q : p) xh::dev_free(c, *q);
            }
        } fr{ctx, {&X, &Y, &M, &W, &Z}};
        alloc(&X, nkk);
        alloc(&Y, nkk);
        alloc(&M, nkk);
        alloc(&W, nkn);
        alloc(&Z, nmk);
        GemmPlan px(ctx, gx, 1, 0), py(ctx, gy, 0, 1), pw(ctx, gw, 0, 0), pu(ctx, gu, 0, 0);
        px.run(U, dU, X, 0.0);
        py.run(V, dV, Y, 0.0);
        SbItem* d_items = xh::dev_upload(ctx, act);
        const unsigned gx_blocks = static_cast<unsigned>(std::min<int64_t>((max_e + SB_THREADS * 4 - 1) / (SB_THREADS * 4), 1024));
        hipLaunchKernelGGL(k_svd_bwd_prep, dim3(std::max(gx_blocks, 1u), static_cast<unsigned>(act.size()), 3), dim3(SB_THREADS), 0, ctx->stream, d_items, S, dS,
                           X, Y, dU, dV, M, W, Z);
        xh::launch_check("k_svd_bwd_prep");
        xh::dev_free(ctx, d_items);
        pw.run(M, V, W, 1.0);
        pu.run(U, W, dA, 1.0);
        pu.run(Z, V, dA, 1.0);
    });
}

// ---------------------------------------------------------------- QR adjoint
// The reference records no adjoint for tQR (include/Xped/AD/ADTensor.hpp has none); this is the
// standard one.  Per sector, A = Q R with k = min(m, n), Q m x k, R k x n upper triangular with a
// nonzero diagonal (full rank).  For m >= n (R square):
//
//   dA += (dQ + Q copyltu(M)) R^-T,   M = R dR^T - dQ^T Q,   copyltu(M)_ij = M_max(i,j),min(i,j),
//
// which follows from dA = dQ' R + Q dR' with Q^T dQ' antisymmetric and dR' upper triangular.
// For m < n, A = Q [R1 R2] splits into the square factorisation X = Q R1 and Y = Q R2: Y's part
// of dA is Q dR2, and X's part is the square formula with dQ + Y dR2^T in place of dQ.
// One workgroup per sector in four phases (products with no data reuse: meant for the small blocks
// of the reference's tQR callers).
namespace {
constexpr int QB_THREADS = 256;
struct QbItem {
    int64_t a_off, q_off, r_off, w_off, m_off;  // w_off: two m x k scratch blocks; m_off: k x k
    int32_t m, n, k, pad;
};
__global__ __launch_bounds__(QB_THREADS) void k_qr_bwd(const QbItem* __restrict__ items, const double* __restrict__ A, const double* __restrict__ Q,
                                                      const double* __restrict__ R, const double* __restrict__ dQ, const double* __restrict__ dR,
                                                      double* __restrict__ dA, double* __restrict__ work)
{
    const QbItem it = items[blockIdx.x];
    const int m = it.m, n = it.n, k = it.k, tid = threadIdx.x;
    const double* a = A + it.a_off;
    const double* q = Q + it.q_off;    // m x k, ld m
    const double* r = R + it.r_off;    // k x n, ld k
    const double* dq = dQ + it.q_off;
    const double* dr = dR + it.r_off;
    double* da = dA + it.a_off;        // m x n, ld m
    double* W1 = work + it.w_off;      // dQ (+ Y dR2^T), m x k
    double* W2 = W1 + int64_t(m) * k;  // dQ' + Q copyltu(M), then its solve, m x k
    double* M = work + it.m_off;       // k x k
    // phase 0: W1 = dQ + Y dR2^T, and dA[:, k:] += Q dR2 (wide blocks only)
    for(int64_t e = tid; e < int64_t(m) * k; e += QB_THREADS) {
        const int i = int(e % m), j = int(e / m);
        double s = dq[e];
        for(int l = k; l < n; ++l) s += a[int64_t(l) * m + i] * dr[int64_t(l) * k + j];
        W1[e] = s;
    }
    for(int64_t e = tid; e < int64_t(m) * (n - k); e += QB_THREADS) {
        const int i = int(e % m), l = k + int(e / m);
        double s = 0.0;
        for(int j = 0; j < k; ++j) s += q[int64_t(j) * m + i] * dr[int64_t(l) * k + j];
        da[int64_t(l) * m + i] += s;
    }
    __syncthreads();
    // phase 1: M = R1 dR1^T - W1^T Q  (R1, dR1: the first k columns)
    for(int64_t e = tid; e < int64_t(k) * k; e += QB_THREADS) {
        const int i = int(e % k), j = int(e / k);
        double s = 0.0;
        for(int c = i; c < k; ++c) s += r[int64_t(c) * k + i] * dr[int64_t(c) * k + j];
        for(int x = 0; x < m; ++x) s -= W1[int64_t(i) * m + x] * q[int64_t(j) * m + x];
        M[e] = s;
    }
    __syncthreads();
    // phase 2: W2 = W1 + Q copyltu(M)
    for(int64_t e = tid; e < int64_t(m) * k; e += QB_THREADS) {
        const int x = int(e % m), j = int(e / m);
        double s = W1[e];
        for(int i = 0; i < k; ++i) s += q[int64_t(i) * m + x] * (i >= j ? M[int64_t(j) * k + i] : M[int64_t(i) * k + j]);
        W2[e] = s;
    }
    __syncthreads();
    // phase 3: row by row, z R1^T = w by back substitution (R1 upper), then dA[:, :k] += z
    for(int x = tid; x < m; x += QB_THREADS) {
        for(int i = k - 1; i >= 0; --i) {
            double s = W2[int64_t(i) * m + x];
            for(int j = i + 1; j < k; ++j) s -= r[int64_t(j) * k + i] * W2[int64_t(j) * m + x];
            W2[int64_t(i) * m + x] = s / r[int64_t(i) * k + i];
        }
        for(int i = 0; i < k; ++i) da[int64_t(i) * m + x] += W2[int64_t(i) * m + x];
    }
}
}  // namespace

extern "C" int xh_qr_bwd(xh_ctx* ctx, const xh_qr_bwd_item* items, int64_t nitems, const double* A, const double* Q, const double* R, const double* dQ,
                         const double* dR, double* dA)
{
    return guard(ctx, [&] {
        if(nitems <= 0) return;
        std::vector<QbItem> act;
        int64_t work = 0;
        for(int64_t g = 0; g < nitems; ++g) {
            const xh_qr_bwd_item& it = items[g];
            if(it.m < 0 || it.n < 0 || it.a_off < 0 || it.q_off < 0 || it.r_off < 0) throw xh::Error(XH_EINVAL, "qr adjoint: bad item");
            if(it.m == 0 || it.n == 0) continue;
            const int32_t k = std::min(it.m, it.n);
            QbItem q{it.a_off, it.q_off, it.r_off, work, 0, it.m, it.n, k, 0};
            work += 2 * int64_t(it.m) * k;
            q.m_off = work;
            work += int64_t(k) * k;
            act.push_back(q);
        }
        if(act.empty()) return;
        double* w = nullptr;
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&w), size_t(work) * sizeof(double)), "hipMallocAsync(qr adjoint)");
        QbItem* d_items = nullptr;
        try {
            d_items = xh::dev_upload(ctx, act);
            hipLaunchKernelGGL(k_qr_bwd, dim3(static_cast<unsigned>(act.size())), dim3(QB_THREADS), 0, ctx->stream, d_items, A, Q, R, dQ, dR, dA, w);
            xh::launch_check("k_qr_bwd");
        } catch(...) {
            xh::dev_free(ctx, d_items);
            xh::dev_free(ctx, w);
            throw;
        }
        xh::dev_free(ctx, d_items);
        xh::dev_free(ctx, w);
    });
}
