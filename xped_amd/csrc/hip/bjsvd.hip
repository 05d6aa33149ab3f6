// FP64 SVD of large sector blocks by parallel one-sided block Jacobi.
//
// Each block A (m x n, worked on as A^T when m < n, so that it is rows x cols with rows >= cols)
// is copied into W and V = I is kept beside it, W = A V throughout.  The columns of W are cut
// into column blocks of width BJ_W; a sweep pairs every two column blocks once (round-robin
// tournament).  A step handles all pairs of all sector blocks at once:
//   G = [W_a W_b]^T [W_a W_b]          one grouped GEMM, one 2w x 2w tile per pair
//   G = J diag J^T                      k_bj_pair: cyclic Jacobi eigen-solve of G in LDS
//   [W_a W_b] <- [W_a W_b] J, same for V   one grouped GEMM that also moves the updated column
//                                       blocks to where the next step wants its pairs adjacent
// A pair whose columns are already orthogonal (|g_ij| <= tol sqrt(g_ii g_jj), the one-sided
// Jacobi test with tol = 8 sqrt(rows) sqrt(2w) eps to sit above the rounding noise of the Gram
// and update GEMMs; columns of norm below rows eps ||A||_F are numerical zeros and exempt) keeps
// J = I; a block is converged after a sweep without a rotation.  Then sigma_j = ||W_j||, U_j = W_j / sigma_j and V are
// read off in descending order of sigma.  The rotations are orthogonal transformations of the
// columns of A itself, so the result has the backward stability of one-sided Jacobi (Demmel &
// Veselic, SIAM J. Matrix Anal. Appl. 13 (1992) 1204; block form: Becka, Oksa & Vajtersic,
// Parallel Comput. 28 (2002) 243).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

#include "device.hpp"

using xh::check;

namespace {

constexpr int BJ_W = 32;          // column block width; pairs are 64 columns wide
constexpr int BJ_THREADS = 256;
constexpr int BJ_MAX_SWEEPS = 60;
constexpr int BJ_INNER_SWEEPS = 2;  // inexact pair solves: the outer sweeps finish the job

struct BjPair {
    int64_t g;       // offset of the pair's n2 x n2 Gram matrix (and of its J)
    int32_t n2;      // width of the pair
    int32_t sector;  // index of the block (for the rotation flag)
    int32_t rows;    // rows of the working matrix (convergence tolerance)
    int32_t pad;
    double z;        // squared norm below which a column counts as numerically zero
};

__device__ __forceinline__ int bj_player(int pos, int step, int np)
{
    if(pos == 0) return 0;
    return ((pos - 1 + step) % (np - 1)) + 1;
}

// Eigen-decomposition G = J diag(g) J^T of one pair's Gram matrix, eigenvalues descending.
__global__ __launch_bounds__(BJ_THREADS) void k_bj_pair(const BjPair* __restrict__ pairs, const double* __restrict__ G, double* __restrict__ J,
                                                         int* __restrict__ flags)
{
    extern __shared__ double sm[];
    __shared__ int need, rotated;
    const BjPair pr = pairs[blockIdx.x];
    const int n = pr.n2;
    double* A = sm;
    double* V = A + n * n;
    double* cs = V + n * n;                          // n doubles: c, s per pair
    int* pi = reinterpret_cast<int*>(cs + n + 2);     // n ints: i, j per pair
    const int tid = threadIdx.x;
    for(int e = tid; e < n * n; e += BJ_THREADS) {
        const int i = e % n, j = e / n;
        A[e] = (i >= j) ? G[pr.g + e] : G[pr.g + j + i * n];
        V[e] = (i == j) ? 1.0 : 0.0;
    }
    if(tid == 0) need = 0;
    __syncthreads();
    // a computed inner product carries rounding noise of about sqrt(rows) eps relative to the
    // column norms, and the GEMM update of a pair adds about sqrt(n2) eps: pairs are orthogonal
    // once below a small multiple of both
    const double tol = 8.0 * fmax(1.0, sqrt(static_cast<double>(pr.rows))) * sqrt(static_cast<double>(n)) * DBL_EPSILON;
    for(int e = tid; e < n * n; e += BJ_THREADS) {
        const int i = e % n, j = e / n;
        if(i < j) {
            const double aii = A[i + i * n], ajj = A[j + j * n], aij = A[i + j * n];
            if(aii > pr.z && ajj > pr.z && fabs(aij) > tol * sqrt(aii) * sqrt(ajj)) need = 1;
        }
    }
    __syncthreads();
    if(need) {
        if(tid == 0) atomicOr(flags + pr.sector, 1);
        const int np = n + (n & 1), npairs = np / 2;
        for(int sweep = 0; sweep < BJ_INNER_SWEEPS; ++sweep) {
            if(tid == 0) rotated = 0;
            __syncthreads();
            for(int step = 0; step < np - 1; ++step) {
                for(int p = tid; p < npairs; p += BJ_THREADS) {
                    int i = bj_player(p, step, np), j = bj_player(np - 1 - p, step, np);
                    if(i > j) {
                        const int t = i;
                        i = j;
                        j = t;
                    }
                    double c = 1.0, s = 0.0;
                    if(j < n) {
                        const double apq = A[i + j * n], app = A[i + i * n], aqq = A[j + j * n];
                        if(fabs(apq) > DBL_EPSILON * sqrt(fabs(app) * fabs(aqq)) && fabs(apq) > 1e-300) {
                            const double theta = (aqq - app) / (2.0 * apq);
                            const double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
                            c = 1.0 / sqrt(t * t + 1.0);
                            s = t * c;
                            rotated = 1;
                        }
                    } else {
                        i = -1;
                    }
                    cs[2 * p] = c;
                    cs[2 * p + 1] = s;
                    pi[2 * p] = i;
                    pi[2 * p + 1] = j;
                }
                __syncthreads();
                // thread (p, q): pair p = tid / G with G = BJ_THREADS / npairs lanes per pair, rows
                // (then columns) q, q + G, ...: no index division inside the loops
                const int G = BJ_THREADS / npairs;
                const int pp = tid / G, q0 = tid - pp * G;
                int ri = -1, rj = 0;
                double rc = 1.0, rs = 0.0;
                if(pp < npairs) {
                    ri = pi[2 * pp];
                    rj = pi[2 * pp + 1];
                    rc = cs[2 * pp];
                    rs = cs[2 * pp + 1];
                }
                const bool act = ri >= 0 && rs != 0.0;
                if(act) {  // columns: A <- A R, V <- V R
                    for(int r = q0; r < n; r += G) {
                        double x = A[r + ri * n], y = A[r + rj * n];
                        A[r + ri * n] = rc * x - rs * y;
                        A[r + rj * n] = rs * x + rc * y;
                        x = V[r + ri * n];
                        y = V[r + rj * n];
                        V[r + ri * n] = rc * x - rs * y;
                        V[r + rj * n] = rs * x + rc * y;
                    }
                }
                __syncthreads();
                if(act) {  // rows: A <- R^T A
                    for(int col = q0; col < n; col += G) {
                        const double x = A[ri + col * n], y = A[rj + col * n];
                        A[ri + col * n] = rc * x - rs * y;
                        A[rj + col * n] = rs * x + rc * y;
                    }
                }
                __syncthreads();
            }
            const int any = rotated;
            __syncthreads();
            if(!any) break;
        }
    }
    // descending eigenvalues (NaN last), ties by index; J columns in that order
    for(int j = tid; j < n; j += BJ_THREADS) {
        const double vj = A[j + j * n] == A[j + j * n] ? A[j + j * n] : -2.0 * DBL_MAX;
        int rank = 0;
        for(int k = 0; k < n; ++k) {
            const double vk = A[k + k * n] == A[k + k * n] ? A[k + k * n] : -2.0 * DBL_MAX;
            rank += (vk > vj) || (vk == vj && k < j);
        }
        pi[rank] = j;
    }
    __syncthreads();
    for(int e = tid; e < n * n; e += BJ_THREADS) {
        const int r = e % n, c = e / n;
        J[pr.g + e] = need ? V[r + pi[c] * n] : (r == c ? 1.0 : 0.0);
    }
}

__global__ void k_set_ones(const int64_t* __restrict__ pos, int64_t n, double* __restrict__ x)
{
    for(int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) x[pos[i]] = 1.0;
}

struct Sec {
    xh_factor_item it;
    bool tr;
    int rows, cols;    // working matrix rows x cols, rows >= cols
    int nb, np, L;     // column blocks, players (even), steps per sweep
    int64_t w, v;      // offsets of W (rows x cols) and V (cols x cols) in the work buffers
    double z = 0.0;    // (rows eps ||A||_F)^2: columns of smaller squared norm are numerical zeros
    int done = 0;      // 0 active, 1 or 2: converged with the result in buffer 0 or 1
    int width(int b) const { return std::min(BJ_W, cols - b * BJ_W); }
    // ordered blocks of step s (pairs adjacent, the dummy player dropped) and pair boundaries
    void layout(int s, std::vector<int>& order, std::vector<std::pair<int, int>>& pairs) const
    {
        order.clear();
        pairs.clear();
        for(int p = 0; p < np / 2; ++p) {
            int a = bj_host_player(p, s), b = bj_host_player(np - 1 - p, s);
            if(a > b) std::swap(a, b);
            if(b >= nb) b = -1;
            pairs.emplace_back(a, b);
            order.push_back(a);
            if(b >= 0) order.push_back(b);
        }
    }
    std::vector<int64_t> col_offsets(int s) const
    {
        std::vector<int> order;
        std::vector<std::pair<int, int>> pairs;
        layout(s, order, pairs);
        std::vector<int64_t> off(nb, 0);
        int64_t c = 0;
        for(int b : order) {
            off[b] = c;
            c += width(b);
        }
        return off;
    }
    int bj_host_player(int pos, int step) const
    {
        if(np == 1) return 0;
        if(pos == 0) return 0;
        return ((pos - 1 + step) % (np - 1)) + 1;
    }
};

struct PlanSet {
    xh_ctx* ctx;
    std::vector<xh_gemm_plan*> gram, update;
    std::vector<int64_t> pair_begin, pair_count;
    int64_t n_pairs = 0;
    BjPair* d_pairs = nullptr;
    ~PlanSet() { clear(); }
    void clear()
    {
        for(auto* p : gram)
            if(p) xh_gemm_plan_destroy(ctx, p);
        for(auto* p : update)
            if(p) xh_gemm_plan_destroy(ctx, p);
        gram.clear();
        update.clear();
        pair_begin.clear();
        pair_count.clear();
        n_pairs = 0;
        xh::dev_free(ctx, d_pairs);
        d_pairs = nullptr;
    }
};

void gemm_plan(xh_ctx* ctx, const std::vector<xh_gemm_item>& items, int ta, int tb, xh_gemm_plan** out)
{
    *out = nullptr;
    if(items.empty()) return;
    const int rc = xh_gemm_plan_create(ctx, items.data(), static_cast<int64_t>(items.size()), ta, tb, out);
    if(rc != XH_OK) throw xh::Error(rc, std::string("block Jacobi plan: ") + xh_last_error(ctx));
}

// Plans for the T steps of one global sweep over the active blocks; block q runs step t mod L_q
// and leaves its columns where its next step wants them.  Returns the Gram buffer size needed.
int64_t build_plans(xh_ctx* ctx, std::vector<Sec>& secs, int T, int64_t nw, PlanSet& ps)
{
    ps.clear();
    // every operand of every item must lie inside its buffer (W/V work buffers of nw doubles,
    // Gram/J buffers of the returned size)
    auto inside = [](int64_t off, int64_t len, int64_t cap) { return off >= 0 && len >= 0 && off + len <= cap; };
    std::vector<BjPair> all_pairs;
    int64_t gsize = 0;
    for(int t = 0; t < T; ++t) {
        std::vector<xh_gemm_item> gram, upd;
        int64_t g = 0;
        ps.pair_begin.push_back(static_cast<int64_t>(all_pairs.size()));
        for(std::size_t q = 0; q < secs.size(); ++q) {
            const Sec& c = secs[q];
            if(c.done) continue;
            const int s = t % c.L, s_next = ((t + 1) % T) % c.L;
            std::vector<int> order;
            std::vector<std::pair<int, int>> pairs;
            c.layout(s, order, pairs);
            const std::vector<int64_t> off = c.col_offsets(s), off_next = c.col_offsets(s_next);
            for(const auto& [a, b] : pairs) {
                const int wa = c.width(a), wb = b >= 0 ? c.width(b) : 0, n2 = wa + wb;
                const int64_t panel_w = c.w + off[a] * c.rows, panel_v = c.v + off[a] * c.cols;
                gram.push_back(xh_gemm_item{panel_w, panel_w, g, n2, n2, c.rows});
                all_pairs.push_back(BjPair{g, n2, static_cast<int32_t>(q), c.rows, 0, c.z});
                upd.push_back(xh_gemm_item{panel_w, g, c.w + off_next[a] * c.rows, c.rows, wa, n2});
                upd.push_back(xh_gemm_item{panel_v, g, c.v + off_next[a] * c.cols, c.cols, wa, n2});
                if(wb) {
                    upd.push_back(xh_gemm_item{panel_w, g + int64_t(wa) * n2, c.w + off_next[b] * c.rows, c.rows, wb, n2});
                    upd.push_back(xh_gemm_item{panel_v, g + int64_t(wa) * n2, c.v + off_next[b] * c.cols, c.cols, wb, n2});
                }
                g += int64_t(n2) * n2;
            }
        }
        ps.pair_count.push_back(static_cast<int64_t>(all_pairs.size()) - ps.pair_begin.back());
        gsize = std::max(gsize, g);
        for(const auto& it : gram)
            if(!inside(it.a_off, int64_t(it.m) * it.k, nw) || !inside(it.c_off, int64_t(it.m) * it.n, g))
                throw xh::Error(XH_EINVAL, "block Jacobi: Gram item out of range");
        for(const auto& it : upd)
            if(!inside(it.a_off, int64_t(it.m) * it.k, nw) || !inside(it.b_off, int64_t(it.k) * it.n, g) || !inside(it.c_off, int64_t(it.m) * it.n, nw))
                throw xh::Error(XH_EINVAL, "block Jacobi: update item out of range");
        xh_gemm_plan *pg = nullptr, *pu = nullptr;
        gemm_plan(ctx, gram, 1, 0, &pg);
        ps.gram.push_back(pg);
        gemm_plan(ctx, upd, 0, 0, &pu);
        ps.update.push_back(pu);
    }
    ps.n_pairs = static_cast<int64_t>(all_pairs.size());
    ps.d_pairs = xh::dev_upload(ctx, all_pairs);
    return gsize;
}

}  // namespace

// Called by xh_block_svd for blocks too large for the single-workgroup Jacobi kernel.
void block_jacobi_svd(xh_ctx* ctx, const std::vector<xh_factor_item>& items, const double* A, double* U, double* S, double* Vt)
{
    if(items.empty()) return;
    std::vector<Sec> secs;
    int64_t nw = 0;
    for(const auto& it : items) {
        Sec c{};
        c.it = it;
        c.tr = it.m < it.n;
        c.rows = std::max(it.m, it.n);
        c.cols = std::min(it.m, it.n);
        c.nb = (c.cols + BJ_W - 1) / BJ_W;
        c.np = c.nb + (c.nb & 1);
        if(c.nb == 1) c.np = 2;
        c.L = c.np - 1;
        c.w = nw;
        nw += int64_t(c.rows) * c.cols;
        c.v = nw;
        nw += int64_t(c.cols) * c.cols;
        secs.push_back(c);
    }
    double *buf[2] = {nullptr, nullptr}, *G = nullptr, *Jm = nullptr;
    int* d_flags = nullptr;
    int64_t* d_ones = nullptr;
    auto cleanup = [&] {
        xh::dev_free(ctx, buf[0]);
        xh::dev_free(ctx, buf[1]);
        xh::dev_free(ctx, G);
        xh::dev_free(ctx, Jm);
        xh::dev_free(ctx, d_flags);
        xh::dev_free(ctx, d_ones);
    };
    PlanSet ps{ctx};
    try {
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&buf[0]), size_t(nw) * sizeof(double)), "hipMallocAsync");
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&buf[1]), size_t(nw) * sizeof(double)), "hipMallocAsync");
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&d_flags), sizeof(int) * secs.size()), "hipMallocAsync");
        // W = A (or A^T) in the layout of step 0, V = the matching column permutation of I
        std::vector<xh_copy_rect> rects;
        std::vector<int64_t> ones;
        for(const auto& c : secs) {
            const std::vector<int64_t> off = c.col_offsets(0);
            for(int b = 0; b < c.nb; ++b) {
                const int wb = c.width(b);
                if(!c.tr)
                    rects.push_back(xh_copy_rect{c.it.a_off + int64_t(b) * BJ_W * c.it.m, c.it.m, c.w + off[b] * c.rows, c.rows, c.rows, wb, 1.0, 0});
                else
                    rects.push_back(xh_copy_rect{c.it.a_off + int64_t(b) * BJ_W, c.it.m, c.w + off[b] * c.rows, c.rows, c.rows, wb, 1.0, 1});
                for(int k = 0; k < wb; ++k) ones.push_back(c.v + (b * BJ_W + k) + (off[b] + k) * int64_t(c.cols));
            }
        }
        {
            std::vector<int64_t> off, len;
            for(const auto& c : secs) {
                off.push_back(c.it.a_off);
                len.push_back(int64_t(c.it.m) * c.it.n);
            }
            std::vector<double> nrm(secs.size());
            const int rc0 = xh_range_sumsq(ctx, off.data(), len.data(), static_cast<int64_t>(off.size()), A, nrm.data());
            if(rc0 != XH_OK) throw xh::Error(rc0, "block Jacobi: block norms");
            for(std::size_t q = 0; q < secs.size(); ++q) {
                if(!std::isfinite(nrm[q])) throw xh::Error(XH_EINVAL, "block SVD: non-finite input block");
                const double f = double(secs[q].rows) * DBL_EPSILON;
                secs[q].z = f * f * nrm[q];
            }
        }
        int rc = xh_copy_rects(ctx, A, buf[0], rects.data(), static_cast<int64_t>(rects.size()));
        if(rc != XH_OK) throw xh::Error(rc, "block Jacobi: initial copy");
        for(const auto& c : secs) check(hipMemsetAsync(buf[0] + c.v, 0, size_t(c.cols) * c.cols * sizeof(double), ctx->stream), "hipMemsetAsync");
        d_ones = xh::dev_upload(ctx, ones);
        hipLaunchKernelGGL(k_set_ones, dim3(static_cast<unsigned>(std::min<int64_t>((ones.size() + 255) / 256, 65535))), dim3(256), 0, ctx->stream, d_ones,
                           static_cast<int64_t>(ones.size()), buf[0]);
        xh::launch_check("k_set_ones");
        const size_t pair_lds = size_t(2 * 2 * BJ_W * 2 * BJ_W + 2 * 2 * BJ_W + 4) * sizeof(double);
        check(hipFuncSetAttribute(reinterpret_cast<const void*>(k_bj_pair), hipFuncAttributeMaxDynamicSharedMemorySize, int(pair_lds)), "hipFuncSetAttribute");
        int cur = 0;
        int64_t gcap = 0;
        bool rebuild = true;
        int T = 0;
        for(int sweep = 0;; ++sweep) {
            if(rebuild) {
                T = 0;
                for(const auto& c : secs)
                    if(!c.done) T = std::max(T, c.L);
                const int64_t gsize = build_plans(ctx, secs, T, nw, ps);
                if(static_cast<int>(ps.gram.size()) != T || static_cast<int>(ps.pair_begin.size()) != T)
                    throw xh::Error(XH_EINVAL, "block Jacobi: inconsistent step plans");
                if(gsize > gcap) {
                    xh::dev_free(ctx, G);
                    xh::dev_free(ctx, Jm);
                    check(xh::dmalloc(ctx, reinterpret_cast<void**>(&G), size_t(gsize) * sizeof(double)), "hipMallocAsync");
                    check(xh::dmalloc(ctx, reinterpret_cast<void**>(&Jm), size_t(gsize) * sizeof(double)), "hipMallocAsync");
                    gcap = gsize;
                }
                rebuild = false;
            }
            check(hipMemsetAsync(d_flags, 0, sizeof(int) * secs.size(), ctx->stream), "hipMemsetAsync");
            for(int t = 0; t < T; ++t) {
                if(ps.gram[t]) {
                    rc = xh_gemm(ctx, ps.gram[t], buf[cur], buf[cur], G, 1.0, 0.0, XH_FP64);
                    if(rc != XH_OK) throw xh::Error(rc, std::string("block Jacobi Gram: ") + xh_last_error(ctx));
                }
                if(ps.pair_begin[t] < 0 || ps.pair_begin[t] + ps.pair_count[t] > ps.n_pairs) throw xh::Error(XH_EINVAL, "block Jacobi: pair range");
                if(ps.pair_count[t]) {
                    hipLaunchKernelGGL(k_bj_pair, dim3(static_cast<unsigned>(ps.pair_count[t])), dim3(BJ_THREADS), pair_lds, ctx->stream,
                                       ps.d_pairs + ps.pair_begin[t], G, Jm, d_flags);
                    xh::launch_check("k_bj_pair");
                }
                if(ps.update[t]) {
                    rc = xh_gemm(ctx, ps.update[t], buf[cur], Jm, buf[1 - cur], 1.0, 0.0, XH_FP64);
                    if(rc != XH_OK) throw xh::Error(rc, std::string("block Jacobi update: ") + xh_last_error(ctx));
                }
                cur = 1 - cur;
            }
            std::vector<int> flags(secs.size());
            check(hipMemcpyAsync(flags.data(), d_flags, sizeof(int) * secs.size(), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            bool active = false;
            for(std::size_t q = 0; q < secs.size(); ++q) {
                if(secs[q].done) continue;
                if(!flags[q]) {
                    secs[q].done = 1 + cur;
                    rebuild = true;
                } else {
                    active = true;
                }
            }
            if(!active) break;
            if(sweep + 1 >= BJ_MAX_SWEEPS) throw xh::Error(XH_ENOTCONV, "block Jacobi SVD did not converge within the sweep limit");
        }
        // singular values = column norms; outputs in descending order
        std::vector<int64_t> roff[2], rlen[2];
        std::vector<std::pair<int, int>> where[2];  // (sector, column)
        for(std::size_t q = 0; q < secs.size(); ++q) {
            const Sec& c = secs[q];
            const int b = c.done - 1;
            for(int j = 0; j < c.cols; ++j) {
                roff[b].push_back(c.w + int64_t(j) * c.rows);
                rlen[b].push_back(c.rows);
                where[b].emplace_back(static_cast<int>(q), j);
            }
        }
        std::vector<std::vector<double>> sig(secs.size());
        for(std::size_t q = 0; q < secs.size(); ++q) sig[q].assign(secs[q].cols, 0.0);
        for(int b = 0; b < 2; ++b) {
            if(roff[b].empty()) continue;
            std::vector<double> nrm(roff[b].size());
            rc = xh_range_sumsq(ctx, roff[b].data(), rlen[b].data(), static_cast<int64_t>(roff[b].size()), buf[b], nrm.data());
            if(rc != XH_OK) throw xh::Error(rc, "block Jacobi: column norms");
            for(std::size_t k = 0; k < nrm.size(); ++k) {
                if(!std::isfinite(nrm[k])) throw xh::Error(XH_EHIP, "block Jacobi SVD: non-finite column norm");
                sig[where[b][k].first][where[b][k].second] = std::sqrt(nrm[k]);
            }
        }
        std::vector<xh_copy_rect> out[2];
        for(std::size_t q = 0; q < secs.size(); ++q) {
            const Sec& c = secs[q];
            const int b = c.done - 1;
            std::vector<int> ord(c.cols);
            std::iota(ord.begin(), ord.end(), 0);
            const std::vector<double>& s = sig[q];
            std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return s[x] > s[y]; });
            std::vector<double> sv(c.cols);
            for(int r = 0; r < c.cols; ++r) {
                const int col = ord[r];
                const double sr = s[col];
                sv[r] = sr;
                const double inv = sr > 0.0 ? 1.0 / sr : 0.0;
                const int64_t wcol = c.w + int64_t(col) * c.rows, vcol = c.v + int64_t(col) * c.cols;
                const int m = c.it.m, n = c.it.n;
                if(!c.tr) {
                    out[b].push_back(xh_copy_rect{wcol, c.rows, c.it.u_off + int64_t(r) * m, m, m, 1, inv, 0});
                    out[b].push_back(xh_copy_rect{vcol, c.cols, c.it.vt_off + r, c.cols, 1, n, 1.0, 1});
                } else {
                    out[b].push_back(xh_copy_rect{vcol, c.cols, c.it.u_off + int64_t(r) * m, m, m, 1, 1.0, 0});
                    out[b].push_back(xh_copy_rect{wcol, c.rows, c.it.vt_off + r, m, 1, n, inv, 1});
                }
            }
            check(hipMemcpyAsync(S + c.it.s_off, sv.data(), sizeof(double) * sv.size(), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        }
        // U and Vt are written from one buffer at a time; both are disjoint outputs
        for(int b = 0; b < 2; ++b) {
            if(out[b].empty()) continue;
            std::vector<xh_copy_rect> ru, rv;
            for(std::size_t k = 0; k < out[b].size(); k += 2) {
                ru.push_back(out[b][k]);
                rv.push_back(out[b][k + 1]);
            }
            rc = xh_copy_rects(ctx, buf[b], U, ru.data(), static_cast<int64_t>(ru.size()));
            if(rc == XH_OK) rc = xh_copy_rects(ctx, buf[b], Vt, rv.data(), static_cast<int64_t>(rv.size()));
            if(rc != XH_OK) throw xh::Error(rc, "block Jacobi: outputs");
        }
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    } catch(...) {
        ps.clear();
        hipStreamSynchronize(ctx->stream);
        cleanup();
        throw;
    }
    ps.clear();
    cleanup();
}
