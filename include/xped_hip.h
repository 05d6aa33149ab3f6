/*
 * xped_hip.h -- device layer of the MI355X block-sparse tensor path (C ABI).
 *
 * Plain pointers and sizes only.  Every function returns 0 (XH_OK) or an error code; the
 * message of the last failure on a context is available from xh_last_error().  One context
 * per host thread; calls on a context are stream-ordered on the context's HIP stream and are
 * not re-entrant.  Plans are immutable after creation and may be shared read-only.
 *
 * The reference performs each of these operations as a loop of per-block CPU library calls
 * behind its compile-time plugin API Xped::PlainInterface (include/Xped/Interfaces/
 * PlainInterface.hpp:102-110).  Each entry point below cites the loop it replaces.
 */
#ifndef XPED_HIP_H_
#define XPED_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    XH_OK = 0,
    XH_EINVAL = 1,   /* invalid argument or plan                                   */
    XH_ENOMEM = 2,   /* device allocation failed (reference: StorageType_contiguous.hpp:91-96 rethrows bad_alloc) */
    XH_EHIP = 3,     /* HIP runtime error                                           */
    XH_ENOTCONV = 4, /* an iterative factorisation did not converge                 */
    XH_ENODEV = 5    /* no usable GPU                                               */
};

/* GEMM precision modes (operands and results FP64 in memory): FP64 matrix cores; FP32; BF16x3
 * (three BF16 products of FP32-rounded operands); XH_OZAKI: FP64 accuracy from the int8 matrix
 * cores by error-free splitting into 8 slices of 7 bits per row of A / column of B (no counterpart
 * in the reference, whose products are FP64 BLAS calls) */
enum { XH_FP64 = 0, XH_FP32 = 1, XH_BF16X3 = 2, XH_OZAKI = 3 };

typedef struct xh_ctx xh_ctx;
typedef struct xh_permute_plan xh_permute_plan;
typedef struct xh_gemm_plan xh_gemm_plan;
typedef struct xh_gather_plan xh_gather_plan;
typedef struct xh_scatter_plan xh_scatter_plan;

/* ---- context and buffers ------------------------------------------------------------ */
int xh_create(int device, xh_ctx** out);
int xh_destroy(xh_ctx* ctx);
const char* xh_last_error(const xh_ctx* ctx);
int xh_device_count(int* count);
int xh_alloc(xh_ctx* ctx, size_t bytes, void** dptr);
int xh_free(xh_ctx* ctx, void* dptr);
/* Device blocks released by xh_free (and the library's own workspaces and plan tables) are kept
 * per context in size classes (four per power of two) and handed out again, in stream order, by
 * later allocations of the same class; idle blocks beyond `bytes` (default 4 GiB; 0 keeps none)
 * go back to the HIP runtime, as do all of them when an allocation fails or the context is
 * destroyed.  No counterpart in the reference, whose storage is host memory
 * (StorageType_contiguous.hpp:91-96). */
int xh_set_device_cache(xh_ctx* ctx, size_t bytes);
/* allocations served by the cache, allocations that reached the runtime, bytes held idle */
int xh_device_cache_stats(xh_ctx* ctx, int64_t* hits, int64_t* misses, size_t* idle_bytes);
int xh_upload(xh_ctx* ctx, void* dst, const void* src, size_t bytes);   /* host -> device, synchronous */
int xh_download(xh_ctx* ctx, void* dst, const void* src, size_t bytes); /* device -> host, synchronous */
int xh_copy(xh_ctx* ctx, void* dst, const void* src, size_t bytes); /* device to device, async on the stream */
int xh_synchronize(xh_ctx* ctx);
void* xh_stream(xh_ctx* ctx); /* the context's hipStream_t */
int xh_fill(xh_ctx* ctx, double* dst, size_t n, double value);
/* dst[off_r .. off_r + len_r) = value for the nranges (off_r, len_r) pairs of ranges[2 nranges],
 * each inside dst[0, n), in one launch: the zero sectors of a product that no GEMM item writes
 * (Tout.setZero() before the products, TensorBase.cpp:256-258) */
int xh_fill_ranges(xh_ctx* ctx, double* dst, size_t n, const int64_t* ranges, int64_t nranges, double value);
/* Standard normal entries from a counter-based generator (splitmix64 + Box-Muller), used for the
 * test matrices of the randomized block SVD; no counterpart in the reference. */
int xh_fill_gaussian(xh_ctx* ctx, double* dst, size_t n, uint64_t seed);
/* out[g] (host) = sum of squares of x[off[g] .. off[g] + len[g]): the per-sector squared
 * Frobenius norms that squaredNorm (TensorBase.cpp:58-64) sums with degeneracy weights. */
int xh_range_sumsq(xh_ctx* ctx, const int64_t* off, const int64_t* len, int64_t nranges, const double* x, double* out);

/* ---- grouped permute ------------------------------------------------------------------
 * Replaces the subblock loop of Tensor::permute_impl (src/Core/Tensor.cpp:83-146, 148-214):
 * for every output subblock, dst = sum_s coeff_s * shuffle_p(src_s), accumulated in list order.
 * Output leg k of a task reads source leg p[k]; its stride inside the source subblock is
 * sstride[k] in the source row index (is_col[k] == 0) or column index (is_col[k] == 1). */
#define XH_MAXLEGS 20
typedef struct {
    int64_t dst_off, dst_ld, rows, cols;
    int32_t nsrc, src_begin;
    uint32_t ext[XH_MAXLEGS];
    int64_t sstride[XH_MAXLEGS];
    uint8_t is_col[XH_MAXLEGS];
} xh_permute_task;
typedef struct {
    int64_t src_off, src_ld;
    double coeff;
} xh_permute_source;
typedef struct {
    int32_t nlegs, new_rank;
    int64_t ntasks, nsources;
    const xh_permute_task* tasks;
    const xh_permute_source* sources;
} xh_permute_desc;
int xh_permute_plan_create(xh_ctx* ctx, const xh_permute_desc* desc, xh_permute_plan** out);
int xh_permute_plan_destroy(xh_ctx* ctx, xh_permute_plan* plan);
int xh_permute(xh_ctx* ctx, const xh_permute_plan* plan, const double* src, double* dst);
/* bytes moved by one xh_permute (8 * (read + written elements)) */
int xh_permute_plan_bytes(const xh_permute_plan* plan, double* bytes);

/* ---- grouped GEMM ---------------------------------------------------------------------
 * Replaces the matched-sector loop of TensorBase::operator* (src/Core/TensorBase.cpp:262-279)
 * calling MatrixInterface::prod (src/Interfaces/MatrixInterface_Eigen_impl.cpp:155-160):
 * C_g = alpha * op(A_g) * op(B_g) + beta * C_g for every item g in one persistent launch.
 * Blocks are column-major; lda = m (or k when transposed), ldb = k (or n), ldc = m. */
typedef struct {
    int64_t a_off, b_off, c_off;
    int32_t m, n, k;
} xh_gemm_item;
int xh_gemm_plan_create(xh_ctx* ctx, const xh_gemm_item* items, int64_t nitems, int trans_a, int trans_b, xh_gemm_plan** out);
int xh_gemm_plan_destroy(xh_ctx* ctx, xh_gemm_plan* plan);
int xh_gemm(xh_ctx* ctx, const xh_gemm_plan* plan, const double* A, const double* B, double* C, double alpha, double beta, int precision);
int xh_gemm_plan_flops(const xh_gemm_plan* plan, double* flops);
/* Device bytes of the split-K partials one xh_gemm call of this plan needs (0: none). */
int xh_gemm_plan_workspace(const xh_gemm_plan* plan, int64_t* bytes);
/* xh_gemm that allocates nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It runs the same kernels as xh_gemm, with the same results bit for bit, but the
 * split-K partials go to a fixed per-device array of XH_GEMM_SCRATCH_BYTES bytes.  The library
 * holds that array on every GPU it runs on until the process ends.  A plan whose partials need more
 * (xh_gemm_plan_workspace) fails with XH_EINVAL.  Every split-K chunk writes its whole partial
 * block before the reduction reads it, so nothing in the array needs to start at zero.  All calls
 * on one GPU share the array: calls that overlap in time on that GPU (several contexts or streams,
 * or graphs replayed concurrently) race and give wrong results.  No timing events and no autotuning
 * are used.  The 64 x 64 tile list that the FP32/BF16x3 modes (and the 64 x 64 FP64 tiling) need is
 * built by a plan's first xh_gemm call in such a mode; without it the call fails with XH_EINVAL. */
#define XH_GEMM_SCRATCH_BYTES ((size_t)32 << 20)
int xh_gemm_noalloc(xh_ctx* ctx, const xh_gemm_plan* plan, const double* A, const double* B, double* C, double alpha, double beta, int precision);
/* ---- permute fused into the grouped GEMM (gather GEMM) -----------------------------------
 * Tensor::contract (Tensor.hpp:328-351) permutes its left operand and then multiplies
 * (TensorBase.cpp:262-279).  xh_gemm_gather computes C = alpha * permute(src) * B + beta * C for
 * every item of a GEMM plan without materialising permute(src): each item's A block is read
 * element by element through the permute's task tables.  No counterpart in the reference.
 *   - perm: the permute of src (as for xh_permute_plan_create); every task must have exactly one
 *     source with coefficient +1 or -1 (abelian symmetries);
 *   - items[j], for GEMM item j: its A block is the permute's output sector whose tasks are
 *     task_base + a * ncod + b (domain tree a < ndom, codomain tree b < ncod); starts[starts_off ..]
 *     holds the ndom + 1 domain-tree row starts, then the ncod + 1 codomain-tree column starts;
 *   - src_elems: size of src in doubles (every source access is checked against it here).
 * Only plans whose every item is a skinny FP64 item with A the long operand (k <= 32, n <= 32, n < m)
 * and no transposes qualify (xh_gemm_gather_eligible); the result equals
 * xh_permute + xh_gemm(XH_FP64) bit for bit. */
typedef struct {
    int32_t task_base, ndom, ncod, pad;
    int64_t starts_off;
} xh_gather_item;
typedef struct {
    xh_permute_desc perm;
    int64_t nitems;
    const xh_gather_item* items;
    int64_t nstarts;
    const int32_t* starts;
    int64_t src_elems;
} xh_gather_desc;
int xh_gemm_gather_eligible(const xh_gemm_plan* plan, int* ok);
int xh_gather_plan_create(xh_ctx* ctx, const xh_gemm_plan* gemm, const xh_gather_desc* desc, xh_gather_plan** out);
int xh_gather_plan_destroy(xh_ctx* ctx, xh_gather_plan* plan);
int xh_gemm_gather(xh_ctx* ctx, const xh_gemm_plan* gemm, const xh_gather_plan* gather, const double* src, int64_t src_elems, const double* B, double* C,
                   double alpha, double beta);
/* The product's own permute folded into the same launch (the result permute of contract): each
 * product element is written straight to its place in permute(product), as 0 + c v with c = +-1,
 * and the product is never materialised.  No counterpart in the reference.
 *   - perm: the permute of the product (output subblocks as tasks, one source each, coefficient
 *     +-1), p its leg permutation (output leg k is product leg p[k]);
 *   - per GEMM item: grid_off, ncod and cstarts_off locate its task grid (task of the product
 *     subblock in domain tree a and codomain tree c: grid[grid_off + a * ncod + c]) and its
 *     ncod + 1 column starts in cstarts; the domain trees are the gather plan's;
 *   - dst_elems: size of the permuted product (every destination access is checked against it).
 * Sectors of the product that no GEMM item writes are not touched: the caller zeroes them.  The
 * result equals xh_gemm_gather (beta = 0) + xh_permute bit for bit. */
typedef struct {
    xh_permute_desc perm;
    const int32_t* p;
    int64_t nitems;
    const int64_t* grid_off;
    const int32_t* ncod;
    const int64_t* cstarts_off;
    int64_t ngrid;
    const int32_t* grid;
    int64_t ncstarts;
    const int32_t* cstarts;
    int64_t dst_elems;
} xh_scatter_desc;
int xh_scatter_plan_create(xh_ctx* ctx, const xh_gather_plan* gather, const xh_scatter_desc* desc, xh_scatter_plan** out);
int xh_scatter_plan_destroy(xh_ctx* ctx, xh_scatter_plan* plan);
int xh_gemm_gather_scatter(xh_ctx* ctx, const xh_gemm_plan* gemm, const xh_gather_plan* gather, const xh_scatter_plan* scatter, const double* src,
                           int64_t src_elems, const double* B, double* dst, int64_t dst_elems, double alpha);
/* GEMM tile autotuner (no counterpart in the reference, whose per-block BLAS calls have no tiles
 * to choose): with enable != 0, the first FP64 call of a plan that has 128 x 128 tiles, with
 * beta = 0, runs both tilings (128 x 128 + 64 x 64, and 64 x 64 only), timed with HIP events
 * (one host synchronisation), and runs the faster last so that C holds exactly its result; the
 * choice is cached per plan signature (transposes, split-K use, log2 buckets of item count, largest
 * m, n, k and total flops) and reused by later plans of the same signature.  Off by default. */
int xh_set_gemm_autotune(xh_ctx* ctx, int enable);
/* autotuner state: cached signatures, tuning runs, cache hits, signatures that chose 64 x 64 only */
int xh_gemm_autotune_stats(xh_ctx* ctx, int64_t* entries, int64_t* tuned, int64_t* hits, int64_t* chose_64);
/* FP64 tiling of sectors with both output sides >= 192: 0 = 128 x 128 tiles (default), 1 = 64 x 64
 * tiles like the other sectors.  A tuning switch; no counterpart in the reference. */
int xh_set_gemm_tiles(xh_ctx* ctx, int mode);
/* HIP-event timing of every xh_gemm launch on the context stream (measurement; no counterpart in
 * the reference): enable (1) / disable (0) / keep (-1); returns the summed kernel time, the flops of
 * the timed launches and their count since the last reset (reset != 0 clears after reading). */
int xh_gemm_timing(xh_ctx* ctx, int enable, int reset, double* kernel_ms, double* flops, int64_t* launches);
/* one timed launch (0 .. launches - 1 since the last reset): its kernel time, the plan's transposes
 * and the caller's tag (trans[3] = {ta, tb, tag}) and the (m, n, k) of its first `cap` items
 * (mnk[3 * cap]); *nitems = all items (measurement, to replay a workload's launches; no
 * counterpart in the reference) */
int xh_gemm_launch_log(xh_ctx* ctx, int64_t launch, double* ms, int32_t* trans, int32_t* mnk, int64_t cap, int64_t* nitems);
/* the tag recorded with the following timed launches (the tensor layer tags its contraction
 * products 1, everything else runs untagged, 0); measurement only */
int xh_set_gemm_tag(xh_ctx* ctx, int tag);
/* HIP-event timing of every xh_permute call, as xh_gemm_timing: summed kernel time, the bytes the
 * timed calls move (8 (elements read + written)) and their count (measurement; no counterpart in
 * the reference) */
int xh_permute_timing(xh_ctx* ctx, int enable, int reset, double* kernel_ms, double* bytes, int64_t* launches);
/* an empty kernel of `id` workgroups (1..4096) on the context stream: a mark in a kernel trace
 * that delimits a region of interest (measurement; no counterpart in the reference) */
int xh_trace_marker(xh_ctx* ctx, int id);

/* ---- batched factorisations -------------------------------------------------------------
 * Replace the per-sector loops of Tensor::full_svd_impl (src/Core/Tensor.cpp:510-519) over
 * PlainInterface::svd (src/Interfaces/PlainInterface_Eigen_impl.cpp:116-127), Tensor::tQR
 * (Tensor.cpp:757-780) over MatrixInterface::qr, and Tensor::teigh (Tensor.cpp:782-803) over
 * MatrixInterface::eigh.  Block g is m x n column-major at a_off; outputs are thin:
 * U (m x r) at u_off, S (r) at s_off, Vt (r x n) at vt_off with r = min(m, n); singular values
 * descending.  eigh: square blocks, eigenvalues ascending, eigenvectors in columns. */
typedef struct {
    int64_t a_off, u_off, s_off, vt_off;
    int32_t m, n;
} xh_factor_item;
int xh_block_svd(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* A, double* U, double* S, double* Vt);
/* QR: Q (m x r) at u_off, R (r x n) at vt_off, Householder, R with the sign convention of
 * Eigen::HouseholderQR (diagonal may be negative). */
int xh_block_qr(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* A, double* Q, double* R);
int xh_block_eigh(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* A, double* W, double* V);
/* Inverse X = R^-1 of the upper Cholesky factor of G + shift_rel[g] * trace(G) * I for square
 * Gram blocks (a_off -> G, u_off -> X, m == n); the step of (shifted) CholeskyQR used by the
 * randomized block SVD.  Numerically dependent columns get a zero row and column in X.  No
 * counterpart in the reference (its SVD is always the full one, Tensor.cpp:494-521). */
int xh_block_chol_inv(xh_ctx* ctx, const xh_factor_item* items, int64_t nitems, const double* G, double* X, const double* shift_rel);

/* ---- elementwise / reductions on a whole tensor buffer or its blocks ---------------------- */
int xh_scale(xh_ctx* ctx, double* x, size_t n, double alpha);                          /* TensorBase.hpp:197-209 */
int xh_axpby(xh_ctx* ctx, double* y, const double* x, size_t n, double a, double b);    /* y = a*x + b*y (CoeffBinaryOp); b == 0: y not read */
int xh_max_abs(xh_ctx* ctx, const double* x, size_t n, double* out);                   /* maxNorm, TensorBase.cpp:48-56 */
/* y = x / maxNorm(x), the normalisation of the reference's CTM moves (CTM.cpp:1372, 1383, ...:
 * `C / C.maxNorm()`), without a host round trip: the norm is reduced and applied on the device.
 * A vanished or non-finite norm copies x unchanged and is counted; xh_normalize_status returns
 * the count since the last nonzero report and the first such norm (synchronous; resets). */
int xh_normalize_max_abs(xh_ctx* ctx, const double* x, double* y, size_t n);
int xh_normalize_status(xh_ctx* ctx, int64_t* count, double* first);
int xh_dot(xh_ctx* ctx, const double* x, const double* y, size_t n, double* out);       /* squaredNorm building block */
/* Per block g (m_g x n_g at off_g): out[g] = trace of the block (TensorBase.cpp:26-46). */
int xh_block_trace(xh_ctx* ctx, const int64_t* off, const int32_t* m, int64_t nblocks, const double* x, double* out);
/* Per row of a set of subblocks: x[rows] *= sign (twist, Tensor.cpp:357-377) expressed as a
 * list of (offset, rows, cols, ld, factor) rectangles. */
typedef struct {
    int64_t off, ld;
    int32_t rows, cols;
    double factor;
} xh_rect;
int xh_scale_rects(xh_ctx* ctx, double* x, const xh_rect* rects, int64_t nrects);
/* diagonal maps of square blocks (DiagCoeffUnaryOp, TensorBase.cpp:191-201): mode 0 = sqrt, 1 = inverse, 2 = inverse sqrt;
 * + 4: the off-diagonal entries are set to zero, as the reference's diagonal expression is a
 * diagonal matrix (PlainInterface diagUnaryFunc: M.diagonal().unaryExpr(f).asDiagonal()) */
int xh_block_diag_map(xh_ctx* ctx, const int64_t* off, const int32_t* m, int64_t nblocks, double* x, int mode);
/* copy rectangles: dst block (rows x cols at dst_off, ld dst_ld) = src block (src_off, src_ld) */
typedef struct {
    int64_t src_off, src_ld, dst_off, dst_ld;
    int32_t rows, cols;
    double factor;
    int32_t transpose; /* 1: dst = factor * src^T (src is cols x rows) */
} xh_copy_rect;
int xh_copy_rects(xh_ctx* ctx, const double* src, double* dst, const xh_copy_rect* rects, int64_t nrects);
/* dst block += factor * op(src block): the scatter of the truncated SVD adjoints into the full ones
 * (cutoff_matrices reverse pass, include/Xped/AD/ADTensor.hpp:357-370).  Rectangles must not overlap. */
int xh_add_rects(xh_ctx* ctx, const double* src, double* dst, const xh_copy_rect* rects, int64_t nrects);

/* ---- reverse mode ----------------------------------------------------------------------
 * The adjoint of the per-sector SVD A = U S V (full_svd_impl reverse pass,
 * include/Xped/AD/ADTensor.hpp:262-320, with the 1e-12 regularised F of :279-282), accumulated
 * into dA for every item.  U and dU blocks are m x k (ld m) at u_off, S and dS k x k (ld k, S
 * diagonal) at s_off, V and dV k x n (ld k) at v_off, dA m x n (ld m) at a_off; k <= min(m, n).
 * The permute adjoint (Tensor::permute_adj, src/Core/Tensor.cpp:258-333) is a permute plan built
 * on the host, and the operator* adjoints (ADTensor.hpp:745-746) are xh_gemm plans with
 * trans_a / trans_b set. */
/* exp(factor * M) of square column-major blocks (TensorBase::mexp, TensorBase.cpp:180-183; Eigen's
 * matrix exponential per block, MatrixInterface_Eigen_impl.hpp:177-180): order m <= XH_MEXP_MAX,
 * block at src + src_off, result at dst + dst_off (src == dst in place is allowed). */
#define XH_MEXP_MAX 64
typedef struct {
    int64_t src_off, dst_off;
    int32_t m, pad;
} xh_mexp_item;
int xh_block_mexp(xh_ctx* ctx, const xh_mexp_item* items, int64_t nitems, const double* src, double* dst, double factor);
/* Diagonal-map reverse passes of square blocks (ADTensor.hpp:560-590; DiagCoeffBinaryOp acts on
 * the diagonal only): mode 0 for y = diag_sqrt(x): dx_ii += dy_ii * 0.5 / y_ii; mode 1 for
 * y = diag_inv(x): dx_ii += -dy_ii * y_ii^2; mode 2 for y = diag_sqrt(x).diag_inv():
 * dx_ii += -dy_ii * y_ii^3 / 2. */
int xh_block_diag_bwd(xh_ctx* ctx, const int64_t* off, const int32_t* m, int64_t nblocks, const double* y, const double* dy, double* dx, int mode);
/* x_g += w_g * identity (rows_g x cols_g blocks): the trace reverse pass with WeightedIdentity
 * (ADTensor.hpp:533-545, Tensor.hpp:238-246). */
int xh_block_add_identity(xh_ctx* ctx, const int64_t* off, const int32_t* rows, const int32_t* cols, const double* w, int64_t nblocks, double* x);
/* max |x_i| and the first index attaining it (maxNorm with maxCoeff's location, ADTensor.hpp:450-476) */
int xh_max_abs_index(xh_ctx* ctx, const double* x, size_t n, double* value, int64_t* index);
/* x[index] += v (the maxNorm reverse pass, ADTensor.hpp:462-471) */
int xh_add_at(xh_ctx* ctx, double* x, int64_t index, double v);
typedef struct {
    int64_t u_off, s_off, v_off, a_off;
    int32_t m, n, k;
} xh_svd_bwd_item;
int xh_svd_bwd(xh_ctx* ctx, const xh_svd_bwd_item* items, int64_t nitems, const double* U, const double* S, const double* V, const double* dU,
               const double* dS, const double* dV, double* dA);
/* The adjoint of the per-sector thin QR A = Q R of xh_block_qr (no counterpart in the reference,
 * whose AD tensor records no tQR): with k = min(m, n), Q m x k (ld m) at q_off, R k x n (ld k,
 * upper triangular, nonzero diagonal) at r_off, the same offsets for dQ and dR, A and dA m x n
 * (ld m) at a_off: dA += (dQ + Q copyltu(R dR^T - dQ^T Q)) R^-T for m >= n; for m < n the
 * square part uses dQ + A[:, k:] dR[:, k:]^T and dA[:, k:] += Q dR[:, k:]. */
typedef struct {
    int64_t a_off, q_off, r_off;
    int32_t m, n;
} xh_qr_bwd_item;
int xh_qr_bwd(xh_ctx* ctx, const xh_qr_bwd_item* items, int64_t nitems, const double* A, const double* Q, const double* R, const double* dQ,
              const double* dR, double* dA);

/* ---- learned rank predictor: a small FP32 network nin -> width -> width -> 1 (ReLU), MSE loss,
 * Adam; hand-written forward/backward kernels.  No counterpart in the reference (its SVD is always
 * the full one); used to choose the per-sector ranks of the randomized block SVD.  X is B x nin
 * row-major, Y, T and grad are device FP32 arrays; parameters: W1 (nin x width), b1, W2
 * (width x width), b2, w3 (width), b3. */
typedef struct xh_mlp xh_mlp;
int xh_mlp_create(xh_ctx* ctx, int nin, int width, uint64_t seed, xh_mlp** out);
int xh_mlp_destroy(xh_ctx* ctx, xh_mlp* net);
int xh_mlp_num_params(const xh_mlp* net, int64_t* n);
int xh_mlp_get_params(xh_ctx* ctx, const xh_mlp* net, float* host);
int xh_mlp_set_params(xh_ctx* ctx, xh_mlp* net, const float* host); /* also resets the Adam moments */
int xh_mlp_forward(xh_ctx* ctx, xh_mlp* net, const float* X, int64_t B, float* Y);
/* gradient of mean((y - T)^2) over the batch into grad (device, num_params floats); *loss (host) if non-null */
int xh_mlp_grad(xh_ctx* ctx, xh_mlp* net, const float* X, const float* T, int64_t B, float* grad, float* loss);
int xh_mlp_adam_step(xh_ctx* ctx, xh_mlp* net, const float* grad, float lr, int64_t t);

/* ---- timing helpers (events on the context stream) --------------------------------------- */
int xh_event_record(xh_ctx* ctx, int slot);
int xh_event_elapsed_ms(xh_ctx* ctx, int slot0, int slot1, float* ms);

/* ---- HIP graphs of grouped launches -----------------------------------------------------
 * Stream capture of the context stream: the grouped launches issued between begin and end
 * (permutes, GEMMs, fills, elementwise kernels on existing buffers) become one graph that
 * xh_graph_launch replays on the stream.  Work that allocates or synchronises cannot be
 * captured: a split-K GEMM plan (it allocates its partials) fails with XH_EINVAL while capturing,
 * as do plan creation and host copies; the capture is then abandoned by xh_graph_end.  Replays
 * read and write the buffers named at capture time.  (No counterpart in the reference, whose
 * CPU path has no launches to batch: the hipGraph-captured steps of a GPU port.) */
typedef struct xh_graph xh_graph;
int xh_graph_begin(xh_ctx* ctx);
int xh_graph_end(xh_ctx* ctx, xh_graph** out);
int xh_graph_launch(xh_ctx* ctx, const xh_graph* g);
int xh_graph_destroy(xh_ctx* ctx, xh_graph* g);
/* nonzero while the context stream is being captured */
int xh_graph_capturing(xh_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* XPED_HIP_H_ */
