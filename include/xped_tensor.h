/*
 * xped_tensor.h -- tensor-level C ABI of the MI355X block-sparse tensor path.
 *
 * This is the drop-in surface for the operations of Xped::Tensor<Scalar=double, Rank, CoRank,
 * Symmetry> (include/Xped/Core/Tensor.hpp:59-444).  Block structure (sectors, fusion trees,
 * offsets) is built on the host in integer arithmetic exactly as the reference builds it; the
 * data lives in GPU memory and every data operation runs as HIP kernels.  A runtime created
 * with device < 0 is structure-only: all structural queries work, data operations return
 * XH_ENODEV.
 *
 * Symmetry kinds (two components at most): 1 = U(1) (SpinU1/ChargeU1), 2 = fermionic U(1)
 * (FChargeU1), 3 = SU(2) (SpinSU2, labels q = 2S+1); e.g. (1,0) = U1<SpinU1>,
 * (2,1) = Combined<U1<FChargeU1>, U1<SpinU1>>, (3,0) = SU2<SpinSU2>.
 *
 * Status codes are those of xped_hip.h.  Handles returned through `out` parameters are owned by
 * the caller and released with the matching *_destroy function.
 */
#ifndef XPED_TENSOR_H_
#define XPED_TENSOR_H_

#include <stddef.h>
#include <stdint.h>

#include "xped_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xt_rt xt_rt;
typedef struct xt_basis xt_basis;
typedef struct xt_tensor xt_tensor;

/* ---- runtime ------------------------------------------------------------------------- */
int xt_rt_create(int device, xt_rt** out);
/* Fails with XH_EINVAL (the runtime stays usable) while tensor or CTM handles made on it are
 * alive: they refer to the runtime, so destroy them first.  Bases do not refer to a runtime. */
int xt_rt_destroy(xt_rt* rt);
const char* xt_last_error(const xt_rt* rt);
xh_ctx* xt_rt_ctx(xt_rt* rt);
/* counters: total GEMM flops and permute bytes issued, number of launches */
int xt_rt_counters(const xt_rt* rt, double* gemm_flops, double* permute_bytes, long* n_gemm, long* n_permute);
int xt_rt_clear_caches(xt_rt* rt);
/* bound of the plan caches (permute and GEMM plans, each): past it the least recently used half is
 * dropped, device tables freed in stream order (no counterpart in the reference, which keeps no
 * device plans); default 2048, at least 8 */
int xt_rt_set_max_plans(xt_rt* rt, int64_t max_plans);
/* contract() with the permute of its left operand read inside the product (xh_gemm_gather)
 * wherever the product qualifies (FP64, every sector block a skinny item with the permuted operand
 * on its long side, single-source +-1 permute, no tape recording): enable = 1 (the default);
 * 2 also writes such a product straight into its own permuted layout when that permute is
 * single-source +-1 (xh_gemm_gather_scatter); 0 materialises every permuted tensor; -1 keeps the
 * setting.  *n_fused (may be null) = contractions run fused so far.  Results are the same bit for
 * bit in every mode; the reference always materialises the permuted operands (Tensor.hpp:328-351). */
int xt_rt_set_fuse_permute(xt_rt* rt, int enable, long* n_fused);
/* current cached permute plans, GEMM plans, and plans dropped so far */
int xt_rt_cache_sizes(const xt_rt* rt, int64_t* permute_plans, int64_t* gemm_plans, int64_t* evicted);
/* host time (seconds) spent building permute and GEMM plans, and how many were built */
int xt_rt_host_stats(const xt_rt* rt, double* permute_plan_s, double* gemm_plan_s, long* n_permute_plans, long* n_gemm_plans);
/* breakdown of that host time (seconds): [0] numeric part of the permute plans on the host, [1]
 * their checks and packing, [2] device permute plans, [3] GEMM plans on the host, [4] device GEMM
 * plans; [5] the number of permute plans whose dimension-free part came from the cache (same leg
 * charges); [6] output structures of the permute plans, [7] their dimension-free parts */
int xt_rt_host_timers(const xt_rt* rt, double* t, int n);
/* GEMM timing: enable = 1 times every grouped GEMM launch of the runtime's context with HIP events
 * (contractions, sketch and Gram products, adjoints) and logs the block shapes of the operator*
 * contractions; 0 stops, -1 leaves the setting; returns the kernel time and flops since the last
 * reset (reset = 1 clears them first) and the number of logged contraction blocks.  Measurement
 * only; no counterpart in the reference. */
int xt_rt_gemm_timing(xt_rt* rt, int enable, int reset, double* kernel_ms, double* flops, long* nlog);
/* an empty marker kernel of `id` workgroups (xh_trace_marker) on the runtime's stream */
int xt_rt_marker(xt_rt* rt, int id);
/* the logged block shapes: 3 ints (m, n, k) per block, up to cap blocks */
int xt_rt_gemm_log(const xt_rt* rt, int32_t* mnk, long cap);

/* ---- bases: Qbasis<Symmetry, depth> (include/Xped/Core/Qbasis.hpp) -------------------- */
int xt_basis_create(int kind0, int kind1, xt_basis** out);                 /* empty depth-1 basis */
int xt_basis_destroy(xt_basis* b);
int xt_basis_copy(const xt_basis* b, xt_basis** out);
int xt_basis_push_back(xt_basis* b, const int* q, size_t dim);             /* Qbasis.cpp:34-73 */
int xt_basis_conj(const xt_basis* b, xt_basis** out);                      /* Qbasis.cpp:418-439 */
int xt_basis_combine(const xt_basis* a, const xt_basis* b, xt_basis** out); /* Qbasis.cpp:441-495 */
int xt_basis_forget_history(const xt_basis* b, xt_basis** out);            /* Qbasis.cpp:410-416 */
int xt_basis_sort(xt_basis* b);                                            /* Qbasis.cpp:306-331 */
int xt_basis_sym(const xt_basis* b, int* kind0, int* kind1);
/* depth, number of sectors, dim(), fullDim(), IS_CONJ() */
int xt_basis_info(const xt_basis* b, int* depth, int* nsec, size_t* dim, size_t* full_dim, int* conj);
/* sector i: q (2 ints), num (offset), inner dim, number of fusion trees */
int xt_basis_entry(const xt_basis* b, int i, int* q, size_t* num, size_t* dim, int* ntrees);
/* tree t of sector i.  rank; unc/inter are 2 ints per q (inter has max(rank-2,0) entries);
 * dims (rank uint32), dual (rank ints), tree dim */
int xt_basis_tree(const xt_basis* b, int i, int t, int* rank, int* unc, int* inter, int* coupled, uint32_t* dims, int* dual, uint64_t* dim);
/* FusionTree::swap / permute (FusionTree.cpp:326-412) applied to tree t of sector i of a coupled
 * basis.  Results: up to `cap` trees given as (unc, inter, coupled) in the layout above and a
 * coefficient each; *n receives the number of results. */
int xt_tree_permute(const xt_basis* b, int i, int t, const int* p, int np, int cap, int* n, int* unc, int* inter, int* coupled, double* coeff);
int xt_tree_swap(const xt_basis* b, int i, int t, int pos, int cap, int* n, int* unc, int* inter, int* coupled, double* coeff);

/* ---- Wigner symbols (arguments are 2j) -------------------------------------------------- */
double xt_wigner_3j(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3);
double xt_wigner_6j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6);
double xt_wigner_9j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6, int tj7, int tj8, int tj9);
/* symmetry coefficients (U1.hpp, SU2.hpp, S1xS2.hpp); q arguments are 2-int arrays */
double xt_coeff(int kind0, int kind1, int which, const int* q1, const int* q2, const int* q3, const int* q4, const int* q5, const int* q6);

/* ---- tensors ------------------------------------------------------------------------- */
/* Tensor(basis_domain, basis_codomain) followed by setZero() */
int xt_tensor_create(xt_rt* rt, int rank, int corank, const xt_basis* const* legs, xt_tensor** out);
int xt_tensor_destroy(xt_tensor* t);
int xt_tensor_copy(const xt_tensor* t, xt_tensor** out); /* new handle sharing the (immutable) data */
int xt_tensor_shape(const xt_tensor* t, int* rank, int* corank, int* nsec, size_t* size);
int xt_tensor_sector(const xt_tensor* t, int i, int* q, size_t* offset, size_t* rows, size_t* cols);
int xt_tensor_leg(const xt_tensor* t, int leg, xt_basis** out);       /* uncoupled leg basis */
int xt_tensor_coupled(const xt_tensor* t, int side, xt_basis** out);  /* side 0: coupledDomain, 1: coupledCodomain */
int xt_tensor_upload(xt_tensor* t, const double* host, size_t n);
int xt_tensor_download(const xt_tensor* t, double* host, size_t n);
const double* xt_tensor_device_ptr(const xt_tensor* t);
/* Tensor::setRandom(engine) with std::mt19937(seed): dist 0 uniform[-1,1), 1 normal(0,1) */
int xt_tensor_set_random(xt_tensor* t, uint32_t seed, int dist);

/* Storage wire format: the fields of the reference's StorageType serialisation
 * (StorageType_contiguous.hpp:191-204: data, sectors, uncoupled and coupled bases, offsets) in one
 * binary layout ("XPEDTNS1", documented in csrc/core/storage.cpp); the reader rebuilds the coupled
 * bases and the layout from the legs and refuses archives that disagree.  A structure-only runtime
 * writes and reads the structure alone.  serialize: *n = the archive's size; the bytes are written
 * when buf is non-null and cap >= *n. */
int xt_tensor_save(const xt_tensor* t, const char* path);
int xt_tensor_load(xt_rt* rt, const char* path, xt_tensor** out);
int xt_tensor_serialize(const xt_tensor* t, void* buf, size_t cap, size_t* n);
int xt_tensor_deserialize(xt_rt* rt, const void* buf, size_t n, xt_tensor** out);

/* ---- operations ------------------------------------------------------------------------ */
int xt_permute(const xt_tensor* t, int shift, const int* p, int np, xt_tensor** out);   /* Tensor.cpp:335-355 */
/* structure-level description of the permute plan: number of output subblocks, sources, and the
 * element count moved; works on a structure-only runtime */
int xt_permute_plan_stats(const xt_tensor* t, int shift, const int* p, int np, int64_t* ntasks, int64_t* nsources, int64_t* elems);
/* dump the plan: per task (dst_off, rows, cols, nsrc) and per source (src_off, coeff) */
int xt_permute_plan_dump(const xt_tensor* t, int shift, const int* p, int np, int64_t* task_dst, int64_t* task_rows, int64_t* task_cols, int32_t* task_nsrc,
                         int64_t* src_off, double* src_coeff);
int xt_mul(const xt_tensor* l, const xt_tensor* r, int precision, xt_tensor** out);   /* TensorBase.cpp:237-288 */
/* operator* into an existing tensor that has the product's structure (no allocation, so it can be
 * captured with xh_graph_begin/xh_graph_end once the product's plan exists); XH_EINVAL otherwise */
int xt_mul_into(const xt_tensor* l, const xt_tensor* r, xt_tensor* out, int precision);
int xt_contract(const xt_tensor* a, const int* a1, int n1, const xt_tensor* b, const int* a2, int n2, int rankres, int precision,
                xt_tensor** out); /* Tensor.hpp:328-351 */
int xt_get_permutations(const int* a1, int n1, int rank1, const int* a2, int n2, int rank2, int rankres, int* p1, int* shift1, int* p2, int* shift2,
                        int* pres, int* shiftres); /* Constfct.hpp:55-109 */
/* The SVspec argument of tSVD: std::map<qarray, VectorType> (Tensor.hpp:353-372), an ordered map
 * from sector to its kept singular values (descending). */
typedef struct xt_spec xt_spec;
int xt_spec_create(xt_spec** out);
int xt_spec_destroy(xt_spec* sp);
int xt_spec_size(const xt_spec* sp, int* nsec);
/* entry i in ascending q: its q (2 ints), the number of values and the first `cap` of them */
int xt_spec_entry(const xt_spec* sp, int i, int* q, int* n, double* vals, int cap);
/* Tensor::tSVD(maxKeep, eps_svd, truncWeight, entropy, SVspec, PRESERVE_MULTIPLETS, RETURN_SPEC)
 * (Tensor.hpp:353-362, Tensor.cpp:478-633): *trunc_weight and *entropy are accumulated into
 * (+=); with return_spec != 0 every kept sector's singular values are inserted into SVspec
 * (an entry already present for a sector is kept, as std::map::insert does). */
int xt_tsvd_spec(const xt_tensor* t, size_t max_keep, double eps_svd, double* trunc_weight, double* entropy, xt_spec* SVspec, int preserve_multiplets,
                 int return_spec, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag);
/* tQR reverse pass, accumulated into dA (structure of A): the QR adjoint per sector
 * (xh_qr_bwd).  The reference's AD tensor records no tQR; this is the adjoint a recorded tQR
 * needs.  Every sector needs the middle dimension min(rows, cols) and a full-rank R. */
int xt_qr_bwd(const xt_tensor* A, const xt_tensor* Q, const xt_tensor* R, const xt_tensor* dQ, const xt_tensor* dR, xt_tensor* dA);
/* Tensor::tSVD (Tensor.cpp:478-633) without SVspec (Tensor.hpp:364-372).  *trunc_weight and
 * *entropy are accumulated into (+=). */
int xt_tsvd(const xt_tensor* t, size_t max_keep, double eps_svd, int preserve_multiplets, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag,
            double* trunc_weight, double* entropy);
/* Randomized truncated block SVD with the tSVD interface and truncation rule (Tensor.cpp:478-633,
 * PRESERVE_MULTIPLETS = false): sector q is sketched with min(ceil(rank_factor * k_q) + oversample,
 * m_q, n_q) columns for the target ranks k_q given as nsec pairs (qs[2 i], qs[2 i + 1]) -> ranks[i]
 * (0 for absent sectors), power_iters subspace iterations in sketch_precision.  *safe is 0 when
 * some sketched sector's smallest Ritz value is not 2x below the smallest kept singular value
 * (a singular value the sketch missed could then belong to the kept set); *residual_ratio is the
 * largest such ratio.  The full tSVD is the reference's only SVD; this entry point has no
 * counterpart there. */
int xt_tsvd_randomized(const xt_tensor* t, size_t max_keep, double eps_svd, const int* qs, const int* ranks, int nsec, int oversample,
                       double rank_factor, int power_iters, int sketch_precision, uint64_t seed, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag,
                       double* trunc_weight, int* safe, double* residual_ratio);
/* The two halves of tSVD, as the AD tensor records them (include/Xped/AD/ADTensor.hpp:221-233):
 * full_svd_impl (Tensor.cpp:494-521): untruncated U (uncoupled domain | middle), Sigma, Vdag
 * with middle dimension min(rows, cols) per sector; and cutoff_matrices (Tensor.cpp:523-633)
 * applied to those factors of t (Sigma's diagonal is the singular-value list). */
int xt_svd_full(const xt_tensor* t, xt_tensor** U, xt_tensor** S, xt_tensor** Vdag);
int xt_cutoff(const xt_tensor* t, const xt_tensor* U, const xt_tensor* S, const xt_tensor* Vdag, size_t max_keep, double eps_svd,
              int preserve_multiplets, xt_tensor** Ut, xt_tensor** St, xt_tensor** Vdagt, double* trunc_weight, double* entropy);

/* ---- reverse mode (include/Xped/AD/ADTensor.hpp) ----------------------------------------
 * Each backward accumulates (+=) into adjoint tensors that have the structures of the values,
 * as the reference's reverse callbacks do. */
/* a zero tensor with the structure of t (an adjoint's initial value) */
int xt_tensor_zeros_like(const xt_tensor* t, xt_tensor** out);
/* Y.permute_adj<shift, p...>() (Tensor.cpp:258-333): for Y = X.permute<shift, p...>(), returns the
 * transpose map applied to Y, with the legs of X; the permute reverse pass (ADTensor.hpp:171-185). */
int xt_permute_adj(const xt_tensor* y, int shift, const int* p, int np, xt_tensor** out);
/* operator* reverse pass (ADTensor.hpp:745-746): dA += dC * B^dagger, dB += A^dagger * dC for
 * C = A * B.  dA or dB may be NULL. */
int xt_mul_bwd(const xt_tensor* A, const xt_tensor* B, const xt_tensor* dC, xt_tensor* dA, xt_tensor* dB, int precision);
/* full_svd_impl reverse pass (ADTensor.hpp:262-320): dA += the SVD adjoint per sector, F with the
 * 1e-12 regularisation of :279-282. */
int xt_svd_bwd(const xt_tensor* U, const xt_tensor* S, const xt_tensor* Vdag, const xt_tensor* dU, const xt_tensor* dS, const xt_tensor* dVdag,
               xt_tensor* dA);
/* cutoff_matrices reverse pass (ADTensor.hpp:357-370): adds the truncated adjoints into the
 * leading columns / corner / rows of the full ones. */
int xt_cutoff_bwd(const xt_tensor* dUt, const xt_tensor* dSt, const xt_tensor* dVdagt, xt_tensor* dU, xt_tensor* dS, xt_tensor* dVdag);
/* diag_sqrt (mode 0) / diag_inv (mode 1) reverse passes (ADTensor.hpp:560-590): dx += ... with
 * y the forward result and dy its adjoint; diagonal entries only. */
int xt_diag_bwd(const xt_tensor* y, const xt_tensor* dy, xt_tensor* dx, int mode);
/* trace reverse pass (ADTensor.hpp:533-545): dx += g * WeightedIdentity (Tensor.hpp:238-246). */
int xt_trace_bwd(xt_tensor* dx, double g);
/* maxNorm reverse pass (ADTensor.hpp:450-476): the first entry of largest modulus of x gets
 * dx += g * sign(x). */
int xt_max_norm_bwd(const xt_tensor* x, xt_tensor* dx, double g);
/* Tensor::tQR(bool RETURN_LQ) (Tensor.hpp:390-391, Tensor.cpp:757-780): Q(uncoupledDomain |
 * middle), R(middle | uncoupledCodomain), middle = the smaller coupled side.  return_lq != 0
 * gives the reference's LQ mode: per sector Q^T and R^T in the same tensors, defined only when
 * every sector block is square (other blocks: XH_EINVAL). */
int xt_tqr(const xt_tensor* t, int return_lq, xt_tensor** Q, xt_tensor** R);
int xt_teigh(const xt_tensor* t, xt_tensor** D, xt_tensor** V); /* Tensor.cpp:782-803 */
int xt_adjoint(const xt_tensor* t, xt_tensor** out);            /* AdjointOp.hpp:25-69 */
int xt_twist(const xt_tensor* t, int leg, xt_tensor** out);     /* Tensor.cpp:357-377 */
int xt_scale(const xt_tensor* t, double alpha, xt_tensor** out);
int xt_axpby(const xt_tensor* x, const xt_tensor* y, double a, double b, xt_tensor** out);
int xt_diag_map(const xt_tensor* t, int mode, xt_tensor** out); /* diag_sqrt (0), diag_inv (1), inverse sqrt (2) */
int xt_trace(const xt_tensor* t, double* out);                  /* TensorBase.cpp:26-37 */
int xt_squared_norm(const xt_tensor* t, double* out);           /* TensorBase.cpp:58-64 */
int xt_max_norm(const xt_tensor* t, double* out);               /* TensorBase.cpp:48-56 */
/* the host truncation rule of cutoff_matrices on explicit (q, sigma) pairs; returns per kept
 * sector the q and the kept count (ascending q) */
int xt_truncation_rule(int kind0, int kind1, const int* qs, const double* svs, int64_t n, size_t max_keep, double eps_svd, int preserve_multiplets,
                       int* nsec, int* kept_q, int* kept_n, double* trunc_weight, double* entropy);

/* ---- CTMRG on a 2x2 unit cell (src/PEPS/CTM.cpp, TRank = 2, CORNER projection) ---------- */
typedef struct xt_ctm xt_ctm;
/* A: four site tensors Tensor<2,3>(left, top | right, bottom, phys) for (x,y) = (0,0),(0,1),(1,0),(1,1) */
int xt_ctm_create(xt_rt* rt, const xt_tensor* const* A, size_t chi, xt_ctm** out);
int xt_ctm_destroy(xt_ctm* c);
/* precision: XH_FP64/XH_FP32/XH_BF16X3 for the contractions; svd_mode 0 = full FP64 block SVD,
 * 1 = randomized block SVD with per-sector target ranks; flags: bit 0 records the full spectra,
 * bit 1 checks every projector factor for non-finite entries, bit 2 prints one line per
 * projector decomposition to stderr, bit 3 makes the learned ranks always sketch their features
 * afresh (by default they reuse the previous decomposition's spectrum of the same site and
 * direction when it covers every sector), bit 4 runs the independent contractions of a move in
 * lockstep with shared GEMM launches (by default they run one at a time) */
int xt_ctm_set_options(xt_ctm* c, int precision, int svd_mode, int oversample, int power_iters, int flags);
int xt_ctm_init(xt_ctm* c);                         /* CTM::init, FROM_A (CTM.cpp:137-240) */
int xt_ctm_grow_all(xt_ctm* c);                     /* CTM::grow_all (CTM.cpp:386-420) */
int xt_ctm_compute_rdm(xt_ctm* c);                  /* CTM::computeRDM (CTM.cpp:422-430) */
int xt_ctm_energy(xt_ctm* c, double Jxy, double Jz, double* E); /* avg + CTMSolver.cpp:60-62 */
/* which: 0..7 = C1..C4, T1..T4; 8 = rho_h, 9 = rho_v; site = x*2 + y */
int xt_ctm_tensor(xt_ctm* c, int which, int site, xt_tensor** out);
/* the environment (C1..C4, T1..T4) of src becomes dst's (the reference's CTM copy `Jim = Jack`,
 * CTMSolver.cpp:91-93); the tensors are shared, not copied */
int xt_ctm_copy_env(xt_ctm* dst, const xt_ctm* src);
/* energy after `steps` more grow_all moves and its gradient with respect to the four site tensors
 * (CTMSolver::solve with AD, CTMSolver.cpp:80-137; each move a checkpoint, CTM.cpp:386-420): the
 * environment on entry is a constant; needs svd_mode 0.  dA[site] are new tensors. */
int xt_ctm_energy_grad(xt_ctm* c, int steps, double Jxy, double Jz, double* E, xt_tensor** dA);
/* decompositions whose learned ranks used the previous spectrum / a fresh sketch as features */
int xt_ctm_predictor_stats(const xt_ctm* c, long* from_previous, long* from_sketch);
int xt_ctm_svd_ms(const xt_ctm* c, double* ms);
/* randomized SVD parameters of the CTM decompositions (see xt_tsvd_randomized) */
int xt_ctm_set_rsvd(xt_ctm* c, int oversample, double rank_factor, int power_iters, int sketch_precision);
/* learned target ranks for the randomized SVDs: network parameters (FP32, the xh_mlp layout with
 * nin = 19 features), hidden width and the rank margin; nparams = 0 removes the predictor.  Without
 * a predictor the ranks are the previous step's kept counts.  No counterpart in the reference. */
int xt_ctm_set_predictor(xt_ctm* c, const float* params, int64_t nparams, int width, double margin);
/* the predictor's 19 input features of one sector (see the network description in the sources):
 * the sector's leading singular values top[0..ntop), min(m, n), the largest sketched value of the
 * tensor and the candidate cut in decades below it */
int xt_rank_features(const double* top, int ntop, int dim, double s_glob, double log_cut, float* out);
/* counters: randomized SVDs run, full-SVD fallbacks taken, largest residual ratio seen */
int xt_ctm_rsvd_stats(const xt_ctm* c, long* n_rsvd, long* n_fallback, double* max_residual_ratio);
/* recorded full spectra: number of records, then per record the sector count and flat arrays */
int xt_ctm_num_spectra(const xt_ctm* c, int* n);
int xt_ctm_spectrum(const xt_ctm* c, int rec, int* site, int* dir, int* nsec, int* qs, int* rows, int* cols, int* kept, int* nsv, double* sv, int cap);
int xt_ctm_clear_spectra(xt_ctm* c);
/* target ranks for the next randomized SVD of (site, dir): nsec pairs (q[2], rank) */
int xt_ctm_set_ranks(xt_ctm* c, int site, int dir, const int* qs, const int* ranks, int nsec);

/* ---- reverse-mode recording (stan-math's reverse-pass callbacks of ADTensor.hpp) ---------- */
/* While a tape records, xt_permute, xt_mul, xt_contract, xt_adjoint, xt_twist, xt_scale,
 * xt_axpby, xt_diag_*, xt_tsvd (full SVD + cutoff), xt_tqr (QR, not LQ) and xt_div_trace push
 * their reverse passes (ADTensor.hpp:138-602, 703-763); adjoints are keyed by the value's storage
 * and accumulate.  Seed with xt_tape_seed (adj(y) += alpha * dy) or xt_tape_seed_trace (the
 * adjoint of y.trace() times g), run xt_tape_backward, read adjoints with xt_tape_adjoint (zeros
 * when none reached x).  One tape per runtime. */
int xt_tape_begin(xt_rt* rt);
int xt_tape_pause(xt_rt* rt, int paused); /* stop (1) or resume (0) recording */
int xt_tape_seed(xt_rt* rt, const xt_tensor* y, const xt_tensor* dy, double alpha);
int xt_tape_seed_trace(xt_rt* rt, const xt_tensor* y, double g);
int xt_tape_backward(xt_rt* rt);
int xt_tape_adjoint(xt_rt* rt, const xt_tensor* x, xt_tensor** out);
int xt_tape_end(xt_rt* rt);
/* x * (1 / s.trace()) with the degeneracy-weighted trace (TensorBase.cpp:26-37), recorded with
 * the adjoints of x and s */
int xt_div_trace(const xt_tensor* x, const xt_tensor* s, xt_tensor** out);

/* ---- simple update on a 2x2 unit cell (src/PEPS/TimePropagator.cpp, nearest-neighbour terms) */
/* TensorBase::mexp (TensorBase.cpp:180-183): exp(factor * block) of every square block (order <= 64) */
int xt_tensor_mexp(const xt_tensor* t, double factor, xt_tensor** out);
typedef struct xt_peps xt_peps;
typedef struct xt_tprop xt_tprop;
/* iPEPS state with bond weights: A as for xt_ctm_create (site = x*2 + y), Gs = A and identity
 * weights (iPEPS::initWeightTensors, iPEPS.cpp:575-591); D is the bond dimension that the
 * update truncates to (iPEPS::D) */
int xt_peps_create(xt_rt* rt, const xt_tensor* const* A, size_t D, xt_peps** out);
int xt_peps_destroy(xt_peps* p);
/* which: 0 = Gs, 1 = whs (bond (x,y)-(x+1,y)), 2 = wvs (bond (x,y-1)-(x,y)), 3 = the weighted site
 * tensors applyWeights(Gs, sqrt weights) (iPEPS::updateAtensors, iPEPS.cpp:593-602) */
int xt_peps_tensor(const xt_peps* p, int which, int site, xt_tensor** out);
/* TimePropagator(H, dt): Hh[x*2+y] and Hv[x*2+y] are the bond Hamiltonians Tensor<2,2>(p, p | p, p)
 * of the bonds (x,y)-(x+1,y) and (x,y)-(x,y+1); the gates are exp(-dt H) (initU, :453-487) */
int xt_tprop_create(xt_rt* rt, const xt_tensor* const* Hh, const xt_tensor* const* Hv, double dt, xt_tprop** out);
int xt_tprop_destroy(xt_tprop* t);
int xt_tprop_step(xt_tprop* t, xt_peps* p); /* t_step, nearest-neighbour branch (TimePropagator.cpp:78-105) */
/* one bond update: t_step_h(x, y) (:113-190) or, with vertical != 0, t_step_v(x, y) (:192-226) */
int xt_tprop_bond(xt_tprop* t, xt_peps* p, int vertical, int x, int y);
/* which: 0 = horizontal gates, 1 = vertical gates, 2 = spectrum_h, 3 = spectrum_v (the last
 * un-normalised weights of each bond) */
int xt_tprop_tensor(const xt_tprop* t, int which, int site, xt_tensor** out);

#ifdef __cplusplus
}
#endif

#endif /* XPED_TENSOR_H_ */
