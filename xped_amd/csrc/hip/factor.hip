// Batched per-sector factorisations in FP64: SVD (one-sided Jacobi), QR (Householder, with the
// reflector convention of Eigen::HouseholderQR) and symmetric eigen (parallel cyclic Jacobi).
//
// One workgroup owns one sector block.  Blocks that fit the LDS budget are factorised entirely in
// LDS; larger blocks use a global-memory working copy with the same algorithm.
//
// The reference calls Eigen per sector (PlainInterface::svd, PlainInterface_Eigen_impl.cpp:116-127;
// MatrixInterface::qr, MatrixInterface_Eigen_impl.cpp:249-255; MatrixInterface::eigh,
// MatrixInterface_Eigen_impl.hpp:242-247), serially over sectors (Tensor.cpp:510-519, 762-777,
// 782-803).
#include <algorithm>
#include <cfloat>
#include <cstring>


#include "device.hpp"

using xh::check;
using xh::guard;

namespace {

constexpr int F_THREADS = 256;
constexpr int MAX_SWEEPS = 60;

__device__ __forceinline__ double group_sum(double v, int g)
{
    for(int o = g >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int rr_player(int pos, int step, int n2)
{
    // circle method: position 0 is fixed, the others rotate by one each step
    if(pos == 0) return 0;
    return ((pos - 1 + step) % (n2 - 1)) + 1;
}

// total orders over doubles for rank-based sorting: NaN below (above) every number
__device__ __forceinline__ double nan_low(double x) { return x == x ? x : -DBL_MAX * 2.0; }
__device__ __forceinline__ double nan_high(double x) { return x == x ? x : DBL_MAX * 2.0; }

struct FactorArgs {
    const xh_factor_item* items;
    const double* A;
    double* U;
    double* S;
    double* Vt;
    double* scratch;
    int64_t scratch_stride;
    int use_lds;
    int* status;
};

// The working set of a workgroup: dynamic LDS, or its slice of the global scratch.  Resolved at
// compile time, so that LDS accesses compile to LDS instructions (a pointer that may point to
// either space compiles to flat loads and stores, which wait on both memory counters).
template <bool LDS>
__device__ __forceinline__ double* work_base(double* smem, const FactorArgs& fa)
{
    if constexpr(LDS)
        return smem;
    else
        return fa.scratch + int64_t(blockIdx.x) * fa.scratch_stride;
}

// Working set of the Jacobi SVD for a block of mm x nn (mm >= nn):
//   W (mm*nn), V (nn*nn), sigma (nn), order (nn ints)
__host__ __device__ inline int64_t svd_work_doubles(int64_t mm, int64_t nn) { return mm * nn + nn * nn + nn + (nn + 1) / 2 + 2; }

template <bool LDS>
__global__ __launch_bounds__(F_THREADS) void k_svd_jacobi(FactorArgs fa)
{
    extern __shared__ double smem[];
    __shared__ int rotated;
    const xh_factor_item it = fa.items[blockIdx.x];
    const int m = it.m, n = it.n;
    if(m == 0 || n == 0) return;
    const bool tr = m < n;
    const int mm = tr ? n : m, nn = tr ? m : n;
    double* W = work_base<LDS>(smem, fa);
    double* V = W + int64_t(mm) * nn;
    double* sig = V + int64_t(nn) * nn;
    int* ord = reinterpret_cast<int*>(sig + nn);
    const int tid = threadIdx.x;
    for(int64_t e = tid; e < int64_t(mm) * nn; e += F_THREADS) {
        const int i = static_cast<int>(e % mm), j = static_cast<int>(e / mm);
        W[e] = tr ? fa.A[it.a_off + j + int64_t(i) * m] : fa.A[it.a_off + i + int64_t(j) * m];
    }
    for(int64_t e = tid; e < int64_t(nn) * nn; e += F_THREADS) V[e] = ((e % nn) == (e / nn)) ? 1.0 : 0.0;
    __syncthreads();

    const int n2 = nn + (nn & 1);
    const int npairs = n2 / 2;
    int g = 64;
    while(g > 8 && g * npairs > F_THREADS) g >>= 1;
    const int ngroups = F_THREADS / g;
    const int grp = tid / g, gl = tid % g;
    const double tol = fmax(1.0, sqrt(static_cast<double>(mm))) * DBL_EPSILON;
    int sweep = 0;
    if(nn > 1) {
        for(sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
            if(tid == 0) rotated = 0;
            __syncthreads();
            for(int step = 0; step < n2 - 1; ++step) {
                for(int pbase = 0; pbase < npairs; pbase += ngroups) {
                    const int p = pbase + grp;
                    int i = -1, j = -1;
                    if(p < npairs) {
                        i = rr_player(p, step, n2);
                        j = rr_player(n2 - 1 - p, step, n2);
                        if(i > j) {
                            int t = i;
                            i = j;
                            j = t;
                        }
                        if(j >= nn) i = -1;
                    }
                    double al = 0, be = 0, ga = 0;
                    if(i >= 0) {
                        const double* wi = W + int64_t(i) * mm;
                        const double* wj = W + int64_t(j) * mm;
                        for(int r = gl; r < mm; r += g) {
                            const double x = wi[r], y = wj[r];
                            al = fma(x, x, al);
                            be = fma(y, y, be);
                            ga = fma(x, y, ga);
                        }
                    }
                    al = group_sum(al, g);
                    be = group_sum(be, g);
                    ga = group_sum(ga, g);
                    if(i >= 0 && al > 0.0 && be > 0.0 && fabs(ga) > tol * sqrt(al) * sqrt(be)) {
                        const double zeta = (be - al) / (2.0 * ga);
                        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        const double c = 1.0 / sqrt(1.0 + t * t);
                        const double s = c * t;
                        double* wi = W + int64_t(i) * mm;
                        double* wj = W + int64_t(j) * mm;
                        for(int r = gl; r < mm; r += g) {
                            const double x = wi[r], y = wj[r];
                            wi[r] = c * x - s * y;
                            wj[r] = s * x + c * y;
                        }
                        double* vi = V + int64_t(i) * nn;
                        double* vj = V + int64_t(j) * nn;
                        for(int r = gl; r < nn; r += g) {
                            const double x = vi[r], y = vj[r];
                            vi[r] = c * x - s * y;
                            vj[r] = s * x + c * y;
                        }
                        if(gl == 0) rotated = 1;
                    }
                }
                __syncthreads();
            }
            const int any = rotated;
            __syncthreads();
            if(!any) break;
        }
    }
    if(sweep >= MAX_SWEEPS && tid == 0) atomicExch(fa.status, XH_ENOTCONV);
    // column norms
    for(int j = tid; j < nn; j += F_THREADS) {
        const double* wj = W + int64_t(j) * mm;
        double s = 0;
        for(int r = 0; r < mm; ++r) s = fma(wj[r], wj[r], s);
        sig[j] = sqrt(s);
    }
    __syncthreads();
    // descending order, ties broken by column index; NaN (from non-finite input) sorts last so
    // that the ranks stay a permutation of 0..nn-1 for any input
    for(int j = tid; j < nn; j += F_THREADS) {
        const double sj = nan_low(sig[j]);
        int rank = 0;
        for(int k = 0; k < nn; ++k) {
            const double sk = nan_low(sig[k]);
            rank += (sk > sj) || (sk == sj && k < j);
        }
        ord[rank] = j;
    }
    __syncthreads();
    for(int r = tid; r < nn; r += F_THREADS) fa.S[it.s_off + r] = sig[ord[r]];
    // U (m x nn) and Vt (nn x n)
    for(int64_t e = tid; e < int64_t(m) * nn; e += F_THREADS) {
        const int i = static_cast<int>(e % m), r = static_cast<int>(e / m);
        const int c = ord[r];
        double v;
        if(!tr) {
            const double s = sig[c];
            v = s > 0.0 ? W[i + int64_t(c) * mm] / s : 0.0;
        } else {
            v = V[i + int64_t(c) * nn];
        }
        fa.U[it.u_off + e] = v;
    }
    for(int64_t e = tid; e < int64_t(nn) * n; e += F_THREADS) {
        const int r = static_cast<int>(e % nn), col = static_cast<int>(e / nn);
        const int c = ord[r];
        double v;
        if(!tr) {
            v = V[col + int64_t(c) * nn];
        } else {
            const double s = sig[c];
            v = s > 0.0 ? W[col + int64_t(c) * mm] / s : 0.0;
        }
        fa.Vt[it.vt_off + e] = v;
    }
}

// Householder QR with Eigen's reflector convention (Eigen/src/Householder/Householder.h
// makeHouseholder: beta = -sign(c0) * ||x||, tau = (beta - c0) / beta, v = [1, x_tail / (c0 - beta)]).
// Q has `q_cols` columns (the tensor's middle dimension; columns beyond m are zero) and R has
// `q_cols` rows (rows beyond min(m, n) are zero).  Item fields: u_off -> Q, vt_off -> R,
// s_off -> q_cols (reused as an integer).
template <bool LDS>
__global__ __launch_bounds__(F_THREADS) void k_qr(FactorArgs fa)
{
    extern __shared__ double smem[];
    __shared__ double red[F_THREADS / 64];
    __shared__ double s_beta, s_tau, s_c0;
    const xh_factor_item it = fa.items[blockIdx.x];
    const int m = it.m, n = it.n;
    const int qc = static_cast<int>(it.s_off);
    if(m == 0 || n == 0) return;
    const int kmax = min(m, n);
    double* Wk = work_base<LDS>(smem, fa);  // m x n working copy
    double* tau = Wk + int64_t(m) * n;                                                      // kmax
    const int tid = threadIdx.x;
    for(int64_t e = tid; e < int64_t(m) * n; e += F_THREADS) Wk[e] = fa.A[it.a_off + e];
    __syncthreads();
    for(int k = 0; k < kmax; ++k) {
        double* col = Wk + int64_t(k) * m;
        double part = 0.0;
        for(int r = k + 1 + tid; r < m; r += F_THREADS) part = fma(col[r], col[r], part);
        for(int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
        if((tid & 63) == 0) red[tid >> 6] = part;
        __syncthreads();
        if(tid == 0) {
            double tail = 0;
            for(int w = 0; w < F_THREADS / 64; ++w) tail += red[w];
            const double c0 = col[k];
            double beta, t;
            if(tail <= DBL_MIN) {
                t = 0.0;
                beta = c0;
            } else {
                beta = sqrt(c0 * c0 + tail);
                if(c0 >= 0.0) beta = -beta;
                t = (beta - c0) / beta;
            }
            s_beta = beta;
            s_tau = t;
            s_c0 = c0;
        }
        __syncthreads();
        const double beta = s_beta, t = s_tau, c0 = s_c0;
        if(t != 0.0) {
            const double inv = 1.0 / (c0 - beta);
            for(int r = k + 1 + tid; r < m; r += F_THREADS) col[r] *= inv;
        }
        __syncthreads();
        if(tid == 0) {
            col[k] = beta;
            tau[k] = t;
        }
        // apply H = I - tau v v^T to the trailing columns, v = [1; col[k+1:]]
        if(t != 0.0) {
            for(int j = k + 1 + tid; j < n; j += F_THREADS) {
                double* cj = Wk + int64_t(j) * m;
                double w = cj[k];
                for(int r = k + 1; r < m; ++r) w = fma(col[r], cj[r], w);
                w *= t;
                cj[k] -= w;
                for(int r = k + 1; r < m; ++r) cj[r] = fma(-w, col[r], cj[r]);
            }
        }
        __syncthreads();
    }
    // R: qc x n, upper triangle of the working copy (rows beyond min(m, n) are zero)
    for(int64_t e = tid; e < int64_t(qc) * n; e += F_THREADS) {
        const int r = static_cast<int>(e % qc), c = static_cast<int>(e / qc);
        fa.Vt[it.vt_off + e] = (r < m && r <= c) ? Wk[r + int64_t(c) * m] : 0.0;
    }
    // Q columns 0..qc-1: Q e_c = H_0 H_1 ... H_{kmax-1} e_c (zero for c >= m)
    for(int c = tid; c < qc; c += F_THREADS) {
        double* qcol = fa.U + it.u_off + int64_t(c) * m;
        for(int r = 0; r < m; ++r) qcol[r] = (r == c) ? 1.0 : 0.0;
        if(c >= m) continue;
        for(int k = kmax - 1; k >= 0; --k) {
            const double t = tau[k];
            if(t == 0.0) continue;
            const double* v = Wk + int64_t(k) * m;
            double w = qcol[k];
            for(int r = k + 1; r < m; ++r) w = fma(v[r], qcol[r], w);
            w *= t;
            qcol[k] -= w;
            for(int r = k + 1; r < m; ++r) qcol[r] = fma(-w, v[r], qcol[r]);
        }
    }
}

// Inverse of the upper Cholesky factor of a Gram block G + shift * trace(G) * I (the building
// block of shifted CholeskyQR, Fukaya, Kannan, Nakatsukasa, Yamamoto & Yanagisawa, SIAM J. Sci.
// Comput. 42 (2020) A477).  Item: a_off -> G (n x n), u_off -> X = R^-1 (n x n, upper), s_off ->
// index of the item's relative shift in fa.S.  A pivot that does not exceed 64 eps times its
// column's original diagonal marks a numerically dependent column: its row and column of X are
// zero, so Y X has a zero column there.
//
// One workgroup per block, one barrier per pivot.  The upper triangle of G (then R) is kept
// packed column by column, (i, j) at j (j + 1) / 2 + i, in LDS when it fits (n <= 199) and in a
// global working copy otherwise; X^T is written straight into the output's strict lower triangle
// (X_ij at U[j + i n]) and reflected at the end; sv[k] = 1 / r_kk (0 for a dependent pivot) is the
// diagonal of X.  Pivot step k runs three disjoint phases:
//   (a) row k-1 of R is scaled by sv[k-1] (the division by r_{k-1,k-1}, deferred one step);
//   (b) right-looking update with the unscaled row k: W_ij -= W_ki W_kj / d_k for k < i <= j,
//       columns paired (j, t-1-j) so that every thread gets the same number of entries;
//   (c) column k-1 of X from the finished rows/columns < k-1, 16 lanes per entry:
//       X_{i,k-1} = -sv[k-1] * sum_{l=i}^{k-2} X_il R_{l,k-1}  (X R = I by columns).
// (a) writes row k-1 (columns >= k), (b) reads row k and writes rows > k, (c) reads rows < k-1 of
// R and columns < k-1 of X and writes column k-1 of X, so one barrier per step suffices.
// Blocks in the global working copy use a panel variant (PANEL = true) of the same steps,
// described at its loop; in LDS the per-pivot barrier, not the trailing update's traffic, bounds
// the time, and the plain loop is as fast.
constexpr int CI_THREADS = 1024;
constexpr int CI_LANES = 16;  // lanes per dot product in (c)
constexpr int CI_PANEL = 8;   // pivots per trailing update of the panel variant
constexpr int CI_MAX_N = 40000;  // packed indices of a block fit in 32 bits

__host__ __device__ inline int64_t chol_work_doubles(int64_t n) { return n * (n + 1) / 2 + 4 * n; }
__device__ __forceinline__ int64_t pk(int i, int j) { return int64_t(j) * (j + 1) / 2 + i; }
__device__ __forceinline__ int pk32(int i, int j) { return ((j * (j + 1)) >> 1) + i; }  // n <= CI_MAX_N

template <bool PANEL>
__global__ __launch_bounds__(CI_THREADS) void k_chol_inv(FactorArgs fa)
{
    extern __shared__ double smem[];
    __shared__ double s_part[CI_THREADS / 64];
    const xh_factor_item it = fa.items[blockIdx.x];
    const int n = it.n;
    if(n == 0) return;
    double* W = work_base<!PANEL>(smem, fa);
    const int64_t np = int64_t(n) * (n + 1) / 2;
    double* rd = W + np;  // shifted original diagonal (dependence threshold)
    double* sv = rd + n;  // 1 / r_kk, or 0
    double* Y = sv + n;   // LDS variant: two columns of X in flight
    double* X = fa.U + it.u_off;
    const int tid = threadIdx.x;
    const double* G = fa.A + it.a_off;
    const int64_t nn = int64_t(n) * n;
    double tr = 0.0;
    for(int64_t e = tid; e < nn; e += CI_THREADS) {
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        if(i <= j) {
            const double g = G[e];
            W[pk(i, j)] = g;
            if(i == j) tr += g;
        }
    }
    for(int o = 32; o > 0; o >>= 1) tr += __shfl_xor(tr, o);
    if((tid & 63) == 0) s_part[tid >> 6] = tr;
    __syncthreads();
    {
        double t = 0.0;
        for(int w = 0; w < CI_THREADS / 64; ++w) t += s_part[w];
        const double shift = fa.S[it.s_off] * t;
        for(int k = tid; k < n; k += CI_THREADS) {
            const double d = W[pk(k, k)] + shift;
            W[pk(k, k)] = d;
            rd[k] = d;
        }
    }
    if constexpr(!PANEL) {
        for(int k = 0; k <= n; ++k) {
            __syncthreads();
            double inv_d = 0.0;
            if(k < n) {
                const double d = W[pk32(k, k)];
                const bool ok = d > 64.0 * DBL_EPSILON * rd[k] && d <= DBL_MAX;
                inv_d = ok ? 1.0 / d : 0.0;
                if(tid == 0) sv[k] = ok ? 1.0 / sqrt(d) : 0.0;  // read from step k + 1 on
            }
            if(k >= 1) {
                // (a) deferred scaling of row k - 1
                const double s = sv[k - 1];
                for(int j = k + tid; j < n; j += CI_THREADS) W[pk32(k - 1, j)] *= s;
            }
            if(k + 1 < n && inv_d != 0.0) {
                // (b) columns a = k+1+pp (rows k+1 .. a) and b = n-1-pp (rows k+1 .. b) as one list
                const int t = n - k - 1;
                const int npairs = (t + 1) / 2;
                const int c2 = npairs <= CI_THREADS ? CI_THREADS / npairs : 1;
                const int p0 = npairs <= CI_THREADS ? tid % npairs : tid;
                const int s0 = npairs <= CI_THREADS ? tid / npairs : 0;
                if(s0 < c2) {
                    for(int pp = p0; pp < npairs; pp += (npairs <= CI_THREADS ? npairs : CI_THREADS)) {
                        const int a = pp, b = t - 1 - pp;
                        const int la = a + 1;
                        const int len = (a == b) ? la : la + b + 1;
                        const double wka = W[pk32(k, k + 1 + a)] * inv_d, wkb = W[pk32(k, k + 1 + b)] * inv_d;
                        // entries s0, s0 + c2, ... of the list: column a's rows, then column b's;
                        // a column's rows are contiguous in the packed triangle
                        const int ca = pk32(k + 1, k + 1 + a), cb = pk32(k + 1, k + 1 + b);
                        int e = s0;
                        for(; e < la; e += c2) W[ca + e] -= W[pk32(k, k + 1 + e)] * wka;
                        for(; e < len; e += c2) W[cb + e - la] -= W[pk32(k, k + 1 + e - la)] * wkb;
                    }
                }
            }
            if(k >= 1) {
                // (c) column j = k - 1 of X into Yc, rows i < j, from the columns of X before it
                // (in place in W, the last one still in Yp) and column j of R; then column j - 1
                // of X (computed last step) goes from Yp into W, over R's column j - 1
                const int j = k - 1;
                const double* Yp = Y + ((k - 1) & 1) * n;
                double* Yc = Y + (k & 1) * n;
                constexpr int per = CI_THREADS / CI_LANES;
                const double sj = sv[j];
                for(int base = 0; base < j; base += per) {
                    const int i = base + tid / CI_LANES, part = tid & (CI_LANES - 1);
                    double acc0 = 0.0, acc1 = 0.0;
                    if(i < j) {
                        // X_il for l < j - 1 in place (its diagonal is sv), X_i,j-1 in Yp.  For
                        // l = j - 1 (column j - 1 of W is being written back this step) the load
                        // reads R's column j instead and its value is not used, so that the select
                        // is a value select, not a branch
                        const double yp = Yp[i];
                        int l = i + part;
                        for(; l + CI_LANES < j; l += 2 * CI_LANES) {
                            const bool last = l + CI_LANES == j - 1;
                            const double x0 = W[pk32(i, l)];
                            const double w1 = W[last ? pk32(l + CI_LANES, j) : pk32(i, l + CI_LANES)];
                            const double x1 = last ? yp : w1;
                            acc0 = fma(x0, W[pk32(l, j)], acc0);
                            acc1 = fma(x1, W[pk32(l + CI_LANES, j)], acc1);
                        }
                        if(l < j) {
                            const bool last = l == j - 1;
                            const double w0 = W[last ? pk32(l, j) : pk32(i, l)];
                            const double x0 = last ? yp : w0;
                            acc0 = fma(x0, W[pk32(l, j)], acc0);
                        }
                    }
                    double acc = acc0 + acc1;
                    for(int o = CI_LANES >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
                    if(i < j && part == 0) Yc[i] = -acc * sj;
                }
                if(tid == 0) Yc[j] = sj;
                if(k >= 2)
                    for(int i = tid; i <= j - 1; i += CI_THREADS) W[pk32(i, j - 1)] = Yp[i];
            }
        }
        // the last column of X into W, then X (upper triangle of W, zeros below) to the output
        __syncthreads();
        {
            const double* Yl = Y + (n & 1) * n;
            for(int i = tid; i < n; i += CI_THREADS) W[pk32(i, n - 1)] = Yl[i];
        }
        __syncthreads();
        for(int64_t e = tid; e < nn; e += CI_THREADS) {
            const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
            X[e] = i <= j ? W[pk32(i, j)] : 0.0;
        }
        return;
    } else {
        // panel variant (blocks in the global working copy): rows of W stay unscaled (row k of R is
        // W's row k times sv[k]) and are never written after their pivot step.  Inside a panel of
        // CI_PANEL pivots, (a) updates the panel's remaining rows by pivot k; at the panel's last
        // pivot, (b) applies the rank-CI_PANEL update of all its pivots to the rows below, so the
        // trailing triangle is read and written once per panel.  (c) is as above with R_lj =
        // sv[l] W_lj.  1 / d_r of an earlier pivot of the panel is taken as sv[r]^2.
        for(int k = 0; k <= n; ++k) {
            __syncthreads();
            if(k < n) {
                const double d = W[pk32(k, k)];
                const bool ok = d > 64.0 * DBL_EPSILON * rd[k] && d <= DBL_MAX;
                const double s = ok ? 1.0 / sqrt(d) : 0.0;
                const double inv_d = s * s;
                if(tid == 0) sv[k] = s;  // read from step k + 1 on
                const int k0 = k - k % CI_PANEL;
                const int kl = min(k0 + CI_PANEL, n) - 1;
                const int t = n - k - 1;
                if(k < kl) {
                    if(inv_d != 0.0) {
                        const int ne = (kl - k) * t;
                        for(int e = tid; e < ne; e += CI_THREADS) {
                            const int i = k + 1 + e / t, j = k + 1 + e % t;
                            if(j >= i) W[pk32(i, j)] -= W[pk32(k, i)] * (W[pk32(k, j)] * inv_d);
                        }
                    }
                } else if(t > 0) {
                    // k == kl < n - 1 implies a full panel: k - k0 + 1 == CI_PANEL
                    const int npairs = (t + 1) / 2;
                    const int c2 = npairs <= CI_THREADS ? CI_THREADS / npairs : 1;
                    const int p0 = npairs <= CI_THREADS ? tid % npairs : tid;
                    const int s0 = npairs <= CI_THREADS ? tid / npairs : 0;
                    if(s0 < c2) {
                        double idr[CI_PANEL];
#pragma unroll
                        for(int r = 0; r < CI_PANEL - 1; ++r) idr[r] = sv[k0 + r] * sv[k0 + r];
                        idr[CI_PANEL - 1] = inv_d;
                        for(int pp = p0; pp < npairs; pp += (npairs <= CI_THREADS ? npairs : CI_THREADS)) {
                            const int a = pp, b = t - 1 - pp;
                            const int la = a + 1;
                            const int len = (a == b) ? la : la + b + 1;
                            const int ca = pk32(0, k + 1 + a), cb = pk32(0, k + 1 + b);
                            double fa[CI_PANEL], fb[CI_PANEL];
#pragma unroll
                            for(int r = 0; r < CI_PANEL; ++r) {
                                fa[r] = W[ca + k0 + r] * idr[r];
                                fb[r] = W[cb + k0 + r] * idr[r];
                            }
                            for(int e = s0; e < len; e += c2) {
                                const bool first = e < la;
                                const int i = k + 1 + (first ? e : e - la);
                                const int bi = pk32(k0, i);
                                const int pos = (first ? ca : cb) + i;
                                double acc = W[pos];
#pragma unroll
                                for(int r = 0; r < CI_PANEL; ++r) acc = fma(-W[bi + r], first ? fa[r] : fb[r], acc);
                                W[pos] = acc;
                            }
                        }
                    }
                }
            }
            if(k >= 2) {
                const int j = k - 1;
                const int cj = pk32(0, j);
                constexpr int per = CI_THREADS / CI_LANES;
                const double sj = sv[j];
                for(int base = 0; base < j; base += per) {
                    const int i = base + tid / CI_LANES, part = tid & (CI_LANES - 1);
                    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
                    if(i < j) {
                        const double* xr = X + int64_t(i) * n;
                        int l = i + part;
                        for(; l + 3 * CI_LANES < j; l += 4 * CI_LANES) {
                            const double xg = xr[l], xs = sv[i];  // both read: selecting a value, not an address space
                            const double x0 = (l == i) ? xs : xg;
                            const double x1 = xr[l + CI_LANES], x2 = xr[l + 2 * CI_LANES], x3 = xr[l + 3 * CI_LANES];
                            acc0 = fma(x0 * sv[l], W[cj + l], acc0);
                            acc1 = fma(x1 * sv[l + CI_LANES], W[cj + l + CI_LANES], acc1);
                            acc2 = fma(x2 * sv[l + 2 * CI_LANES], W[cj + l + 2 * CI_LANES], acc2);
                            acc3 = fma(x3 * sv[l + 3 * CI_LANES], W[cj + l + 3 * CI_LANES], acc3);
                        }
                        for(; l < j; l += CI_LANES) {
                            const double xg = xr[l], xs = sv[i];  // both read: selecting a value, not an address space
                            const double x0 = (l == i) ? xs : xg;
                            acc0 = fma(x0 * sv[l], W[cj + l], acc0);
                        }
                    }
                    double acc = (acc0 + acc1) + (acc2 + acc3);
                    for(int o = CI_LANES >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
                    if(i < j && part == 0) X[j + int64_t(i) * n] = -acc * sj;
                }
            }
        }
    }
    __syncthreads();
    // reflect X^T into the upper triangle, zero the lower one, set the diagonal
    for(int64_t e = tid; e < nn; e += CI_THREADS) {
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        if(i < j) {
            const int64_t lo = j + int64_t(i) * n;
            X[e] = X[lo];
            X[lo] = 0.0;
        } else if(i == j) {
            X[e] = sv[j];
        }
    }
}

// Symmetric eigen-decomposition by parallel cyclic Jacobi.  Item: a_off (n x n), s_off -> eigenvalues,
// u_off -> eigenvectors (columns), ascending order.
template <bool LDS>
__global__ __launch_bounds__(F_THREADS) void k_eigh(FactorArgs fa)
{
    extern __shared__ double smem[];
    __shared__ int rotated;
    const xh_factor_item it = fa.items[blockIdx.x];
    const int n = it.m;
    if(n == 0) return;
    double* Am = work_base<LDS>(smem, fa);
    double* V = Am + int64_t(n) * n;
    double* cs = V + int64_t(n) * n;  // per pair: c, s  (n entries)
    int* pr = reinterpret_cast<int*>(cs + n + 1);  // per pair: p, q (n ints)
    double* ev = cs;  // reused after the sweeps
    int* ord = pr;
    const int tid = threadIdx.x;
    for(int64_t e = tid; e < int64_t(n) * n; e += F_THREADS) {
        // symmetrise from the lower triangle as SelfAdjointEigenSolver reads it
        const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
        Am[e] = (i >= j) ? fa.A[it.a_off + e] : fa.A[it.a_off + j + int64_t(i) * n];
        V[e] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int n2 = n + (n & 1), npairs = n2 / 2;
    int sweep = 0;
    if(n > 1) {
        for(sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
            if(tid == 0) rotated = 0;
            __syncthreads();
            for(int step = 0; step < n2 - 1; ++step) {
                for(int p = tid; p < npairs; p += F_THREADS) {
                    int i = rr_player(p, step, n2), j = rr_player(n2 - 1 - p, step, n2);
                    if(i > j) {
                        int t = i;
                        i = j;
                        j = t;
                    }
                    double c = 1.0, s = 0.0;
                    if(j < n) {
                        const double apq = Am[i + int64_t(j) * n];
                        const double app = Am[i + int64_t(i) * n], aqq = Am[j + int64_t(j) * n];
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
                    pr[2 * p] = i;
                    pr[2 * p + 1] = j;
                }
                __syncthreads();
                // columns: A <- A J, V <- V J
                for(int64_t e = tid; e < int64_t(npairs) * n; e += F_THREADS) {
                    const int p = static_cast<int>(e / n), r = static_cast<int>(e % n);
                    const int i = pr[2 * p], j = pr[2 * p + 1];
                    if(i < 0) continue;
                    const double c = cs[2 * p], s = cs[2 * p + 1];
                    if(s == 0.0) continue;
                    double x = Am[r + int64_t(i) * n], y = Am[r + int64_t(j) * n];
                    Am[r + int64_t(i) * n] = c * x - s * y;
                    Am[r + int64_t(j) * n] = s * x + c * y;
                    x = V[r + int64_t(i) * n];
                    y = V[r + int64_t(j) * n];
                    V[r + int64_t(i) * n] = c * x - s * y;
                    V[r + int64_t(j) * n] = s * x + c * y;
                }
                __syncthreads();
                // rows: A <- J^T A
                for(int64_t e = tid; e < int64_t(npairs) * n; e += F_THREADS) {
                    const int p = static_cast<int>(e / n), col = static_cast<int>(e % n);
                    const int i = pr[2 * p], j = pr[2 * p + 1];
                    if(i < 0) continue;
                    const double c = cs[2 * p], s = cs[2 * p + 1];
                    if(s == 0.0) continue;
                    const double x = Am[i + int64_t(col) * n], y = Am[j + int64_t(col) * n];
                    Am[i + int64_t(col) * n] = c * x - s * y;
                    Am[j + int64_t(col) * n] = s * x + c * y;
                }
                __syncthreads();
            }
            const int any = rotated;
            __syncthreads();
            if(!any) break;
        }
    }
    if(sweep >= MAX_SWEEPS && tid == 0) atomicExch(fa.status, XH_ENOTCONV);
    for(int j = tid; j < n; j += F_THREADS) ev[j] = Am[j + int64_t(j) * n];
    __syncthreads();
    for(int j = tid; j < n; j += F_THREADS) {
        const double vj = nan_high(ev[j]);
        int rank = 0;
        for(int k = 0; k < n; ++k) rank += (nan_high(ev[k]) < vj) || (nan_high(ev[k]) == vj && k < j);
        ord[rank] = j;
    }
    __syncthreads();
    for(int r = tid; r < n; r += F_THREADS) fa.S[it.s_off + r] = ev[ord[r]];
    for(int64_t e = tid; e < int64_t(n) * n; e += F_THREADS) {
        const int i = static_cast<int>(e % n), r = static_cast<int>(e / n);
        fa.U[it.u_off + e] = V[i + int64_t(ord[r]) * n];
    }
}

struct Launch {
    std::vector<xh_factor_item> lds_items, glb_items;
    int64_t lds_bytes = 0, glb_stride = 0;
};

size_t lds_budget(xh_ctx* ctx)
{
    static size_t cached = 0;
    if(cached == 0) {
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
        cached = prop.sharedMemPerBlock;
        if(cached > 160 * 1024) cached = 160 * 1024;
    }
    return cached - 1024;  // leave room for the static shared variables
}

template <typename Need>
void run_factor(xh_ctx* ctx, void (*kern_lds)(FactorArgs), void (*kern_global)(FactorArgs), const std::vector<xh_factor_item>& items, Need need_doubles,
                const double* A, double* U, double* S, double* Vt)
{
    const size_t budget = lds_budget(ctx);
    std::vector<xh_factor_item> small, big;
    int64_t small_max = 0, big_max = 0;
    for(const auto& it : items) {
        if(it.m < 0 || it.n < 0) throw xh::Error(XH_EINVAL, "factor: negative extent");
        if(it.m == 0 || it.n == 0) continue;
        const int64_t nd = need_doubles(it);
        if(size_t(nd) * sizeof(double) <= budget) {
            small.push_back(it);
            small_max = std::max(small_max, nd);
        } else {
            big.push_back(it);
            big_max = std::max(big_max, nd);
        }
    }
    int* d_status = nullptr;
    check(xh::dmalloc(ctx, reinterpret_cast<void**>(&d_status), sizeof(int)), "hipMallocAsync");
    check(hipMemsetAsync(d_status, 0, sizeof(int), ctx->stream), "hipMemsetAsync");
    auto go = [&](const std::vector<xh_factor_item>& v, bool lds, int64_t maxd) {
        if(v.empty()) return;
        xh_factor_item* d_items = xh::dev_upload(ctx, v);
        double* scratch = nullptr;
        if(!lds) check(xh::dmalloc(ctx, reinterpret_cast<void**>(&scratch), size_t(maxd) * v.size() * sizeof(double)), "hipMallocAsync");
        FactorArgs fa{d_items, A, U, S, Vt, scratch, maxd, lds ? 1 : 0, d_status};
        auto* kern = lds ? kern_lds : kern_global;
        const size_t shmem = lds ? size_t(maxd) * sizeof(double) : 0;
        if(lds && shmem > 64 * 1024) check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)), "hipFuncSetAttribute");
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(v.size())), dim3(F_THREADS), shmem, ctx->stream, fa);
        xh::launch_check("factor kernel");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        xh::dev_free(ctx, d_items);
        xh::dev_free(ctx, scratch);
    };
    go(small, true, small_max);
    go(big, false, big_max);
    int status = 0;
    check(hipMemcpyAsync(&status, d_status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    xh::dev_free(ctx, d_status);
    if(status != 0) throw xh::Error(XH_ENOTCONV, "Jacobi iteration did not converge within the sweep limit");
}

}  // namespace

// blocks with more columns than this (after transposition) use the block Jacobi path
constexpr int BJ_MIN_COLS = 64;

extern "C" int xh_block_svd(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* A, double* U, double* S, double* Vt)
{
    return guard(ctx, [&] {
        // small blocks: one workgroup each, in LDS; larger ones: block Jacobi over grouped GEMMs
        std::vector<xh_factor_item> v, big;
        const size_t budget = lds_budget(ctx);
        for(int64_t i = 0; i < nitems; ++i) {
            const xh_factor_item& it = items[i];
            if(it.m <= 0 || it.n <= 0) continue;
            const int mn = std::min(it.m, it.n), mx = std::max(it.m, it.n);
            if(mn > BJ_MIN_COLS || size_t(svd_work_doubles(mx, mn)) * sizeof(double) > budget)
                big.push_back(it);
            else
                v.push_back(it);
        }
        block_jacobi_svd(ctx, big, A, U, S, Vt);
        run_factor(
            ctx, k_svd_jacobi<true>, k_svd_jacobi<false>, v,
            [](const xh_factor_item& it) {
                const int64_t mm = std::max(it.m, it.n), nn = std::min(it.m, it.n);
                return svd_work_doubles(mm, nn);
            },
            A, U, S, Vt);
    });
}

extern "C" int xh_block_qr(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* A, double* Q, double* R)
{
    return guard(ctx, [&] {
        std::vector<xh_factor_item> v(items, items + nitems);
        run_factor(
            ctx, k_qr<true>, k_qr<false>, v, [](const xh_factor_item& it) { return int64_t(it.m) * it.n + std::min(it.m, it.n) + 1; }, A, Q, nullptr, R);
    });
}

extern "C" int xh_block_eigh(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* A, double* W, double* V)
{
    return guard(ctx, [&] {
        std::vector<xh_factor_item> v(items, items + nitems);
        for(auto& it : v)
            if(it.m != it.n) throw xh::Error(XH_EINVAL, "eigh: blocks must be square");
        run_factor(
            ctx, k_eigh<true>, k_eigh<false>, v, [](const xh_factor_item& it) { return 2 * int64_t(it.m) * it.m + 2 * int64_t(it.m) + 4; }, A, V, W, nullptr);
    });
}

extern "C" int xh_block_chol_inv(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* G, double* X, const double* shift_rel)
{
    // asynchronous: no status to read back, so no host synchronisation
    return guard(ctx, [&] {
        const size_t budget = lds_budget(ctx);
        std::vector<xh_factor_item> small, big;
        std::vector<double> sh;
        int64_t small_max = 0, big_max = 0;
        for(int64_t i = 0; i < nitems; ++i) {
            xh_factor_item it = items[i];
            if(it.m != it.n) throw xh::Error(XH_EINVAL, "chol_inv: blocks must be square");
            if(it.n < 0 || it.a_off < 0 || it.u_off < 0) throw xh::Error(XH_EINVAL, "chol_inv: negative extent or offset");
            if(it.n > CI_MAX_N) throw xh::Error(XH_EINVAL, "chol_inv: block larger than 40000");
            if(it.n == 0) continue;
            it.s_off = static_cast<int64_t>(sh.size());
            sh.push_back(shift_rel ? shift_rel[i] : 0.0);
            const int64_t need = chol_work_doubles(it.n);
            if(size_t(need) * sizeof(double) <= budget) {
                small.push_back(it);
                small_max = std::max(small_max, need);
            } else {
                big.push_back(it);
                big_max = std::max(big_max, need);
            }
        }
        if(sh.empty()) return;
        xh::PackedUpload up;
        const size_t o_small = up.add(small), o_big = up.add(big), o_sh = up.add(sh);
        up.upload(ctx);
        double* d_sh = up.ptr<double>(o_sh);
        if(!small.empty()) {
            const size_t shmem = size_t(small_max) * sizeof(double);
            if(shmem > 64 * 1024)
                check(hipFuncSetAttribute(reinterpret_cast<const void*>(k_chol_inv<false>), hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)), "hipFuncSetAttribute");
            FactorArgs fa{up.ptr<xh_factor_item>(o_small), G, X, d_sh, nullptr, nullptr, 0, 1, nullptr};
            hipLaunchKernelGGL(k_chol_inv<false>, dim3(static_cast<unsigned>(small.size())), dim3(CI_THREADS), shmem, ctx->stream, fa);
            xh::launch_check("k_chol_inv");
        }
        if(!big.empty()) {
            double* scratch = nullptr;
            check(xh::dmalloc(ctx, reinterpret_cast<void**>(&scratch), size_t(big_max) * big.size() * sizeof(double)), "hipMallocAsync");
            FactorArgs fa{up.ptr<xh_factor_item>(o_big), G, X, d_sh, nullptr, scratch, big_max, 0, nullptr};
            hipLaunchKernelGGL(k_chol_inv<true>, dim3(static_cast<unsigned>(big.size())), dim3(CI_THREADS), 0, ctx->stream, fa);
            xh::launch_check("k_chol_inv");
            xh::dev_free(ctx, scratch);
        }
        up.release_deferred(ctx);
    });
}

