// Reverse-pass operations of the AD tensor (include/Xped/AD/ADTensor.hpp): the adjoints of
// operator*, of full_svd_impl and of cutoff_matrices, on device.  The permute adjoint is
// permute_adj (tensor.cpp).  Every function accumulates into the adjoint tensors it is given,
// as the reference's reverse callbacks do (`curr.adj() += ...`).
#include "ad.hpp"

#include <algorithm>
#include <cmath>

namespace xa {

namespace {

void dcheck(xh_ctx* ctx, int rc, const char* what)
{
    if(rc != XH_OK) throw Error(rc, std::string(what) + ": " + xh_last_error(ctx));
}

void require_same_blocks(const TensorStructure& a, const TensorStructure& b, const char* what)
{
    if(a.sector != b.sector || a.rows != b.rows || a.cols != b.cols || a.size != b.size)
        throw Error(XH_EINVAL, std::string(what) + ": adjoint does not have the structure of its value");
}

}  // namespace

// ADTensor.hpp:745-746 (and the one-sided variants :711, :733):
//   right.adj() += left^dagger * res.adj(),  left.adj() += res.adj() * right^dagger.
// Real blocks: per sector q of A (dom | mid) and B (mid | cod) matched with dC (dom | cod),
//   dA_q += dC_q B_q^T  (a GEMM with trans_b),  dB_q += A_q^T dC_q  (a GEMM with trans_a).
void mul_bwd(const DTensor& A, const DTensor& B, const DTensor& dC, DTensor* dA, DTensor* dB, int precision)
{
    A.rt->require_device();
    const TensorStructure &a = *A.st, &b = *B.st, &c = *dC.st;
    if(a.sym.kind != b.sym.kind || a.sym.kind != c.sym.kind || a.sym.nq != c.sym.nq) throw Error(XH_EINVAL, "mul adjoint: symmetry mismatch");
    if(dA) require_same_blocks(a, *dA->st, "mul adjoint (left)");
    if(dB) require_same_blocks(b, *dB->st, "mul adjoint (right)");
    std::vector<xh_gemm_item> ia, ib;
    for(std::size_t i = 0; i < c.sector.size(); ++i) {
        const Q& q = c.sector[i];
        const int ja = a.find_sector(q), jb = b.find_sector(q);
        if(ja < 0 || jb < 0) continue;
        if(a.rows[ja] != c.rows[i] || b.cols[jb] != c.cols[i] || a.cols[ja] != b.rows[jb]) throw Error(XH_EINVAL, "mul adjoint: block shapes do not match");
        const auto m = static_cast<int32_t>(a.rows[ja]), mid = static_cast<int32_t>(a.cols[ja]), n = static_cast<int32_t>(c.cols[i]);
        if(m == 0 || mid == 0 || n == 0) continue;
        // dA_q (m x mid) += dC_q (m x n) * B_q^T, B_q stored mid x n
        ia.push_back(xh_gemm_item{static_cast<int64_t>(c.offsets[i]), static_cast<int64_t>(b.offsets[jb]), static_cast<int64_t>(a.offsets[ja]), m, mid, n});
        // dB_q (mid x n) += A_q^T * dC_q, A_q stored m x mid
        ib.push_back(xh_gemm_item{static_cast<int64_t>(a.offsets[ja]), static_cast<int64_t>(c.offsets[i]), static_cast<int64_t>(b.offsets[jb]), mid, n, m});
    }
    auto run = [&](const std::vector<xh_gemm_item>& items, int ta, int tb, const double* x, const double* y, double* out, const char* what) {
        if(items.empty()) return;
        xh_gemm_plan* p = nullptr;
        dcheck(A.rt->ctx, xh_gemm_plan_create(A.rt->ctx, items.data(), static_cast<int64_t>(items.size()), ta, tb, &p), what);
        int rc = xh_gemm(A.rt->ctx, p, x, y, out, 1.0, 1.0, precision);
        xh_gemm_plan_destroy(A.rt->ctx, p);
        dcheck(A.rt->ctx, rc, what);
        double f = 0.0;
        for(const auto& g : items) f += 2.0 * g.m * double(g.n) * g.k;
        A.rt->gemm_flops += f;
        A.rt->n_gemm++;
    };
    if(dA) run(ia, 0, 1, dC.data(), B.data(), dA->data(), "mul adjoint (left)");
    if(dB) run(ib, 1, 0, A.data(), dC.data(), dB->data(), "mul adjoint (right)");
}

// full_svd_impl reverse pass (ADTensor.hpp:262-320): for every sector of A whose q is a sector of
// S, dA_q += the SVD adjoint of (U_q, S_q, V_q) with incoming (dU_q, dS_q, dV_q).
void svd_bwd(const DTensor& U, const DTensor& S, const DTensor& Vdag, const DTensor& dU, const DTensor& dS, const DTensor& dVdag, DTensor& dA)
{
    U.rt->require_device();
    require_same_blocks(*U.st, *dU.st, "svd adjoint (U)");
    require_same_blocks(*S.st, *dS.st, "svd adjoint (S)");
    require_same_blocks(*Vdag.st, *dVdag.st, "svd adjoint (Vdag)");
    const TensorStructure &a = *dA.st, &u = *U.st, &s = *S.st, &v = *Vdag.st;
    std::vector<xh_svd_bwd_item> items;
    for(std::size_t i = 0; i < a.sector.size(); ++i) {
        const Q& q = a.sector[i];
        const int js = s.find_sector(q);
        if(js < 0) continue;  // ADTensor.hpp:265-266
        const int ju = u.find_sector(q), jv = v.find_sector(q);
        if(ju < 0 || jv < 0) throw Error(XH_EINVAL, "svd adjoint: U or Vdag lacks a sector of Sigma");
        const std::size_t k = s.rows[js];
        if(s.cols[js] != k || u.rows[ju] != a.rows[i] || u.cols[ju] != k || v.rows[jv] != k || v.cols[jv] != a.cols[i] || k > std::min(a.rows[i], a.cols[i]))
            throw Error(XH_EINVAL, "svd adjoint: block shapes do not match");
        items.push_back(xh_svd_bwd_item{static_cast<int64_t>(u.offsets[ju]), static_cast<int64_t>(s.offsets[js]), static_cast<int64_t>(v.offsets[jv]),
                                        static_cast<int64_t>(a.offsets[i]), static_cast<int32_t>(a.rows[i]), static_cast<int32_t>(a.cols[i]),
                                        static_cast<int32_t>(k)});
    }
    if(items.empty()) return;
    dcheck(U.rt->ctx,
           xh_svd_bwd(U.rt->ctx, items.data(), static_cast<int64_t>(items.size()), U.data(), S.data(), Vdag.data(), dU.data(), dS.data(), dVdag.data(), dA.data()),
           "svd adjoint");
}

void qr_bwd(const DTensor& A, const DTensor& Q, const DTensor& R, const DTensor& dQ, const DTensor& dR, DTensor& dA)
{
    A.rt->require_device();
    require_same_blocks(*Q.st, *dQ.st, "qr adjoint (Q)");
    require_same_blocks(*R.st, *dR.st, "qr adjoint (R)");
    require_same_blocks(*A.st, *dA.st, "qr adjoint (A)");
    const TensorStructure &a = *A.st, &q = *Q.st, &r = *R.st;
    std::vector<xh_qr_bwd_item> items;
    for(std::size_t i = 0; i < a.sector.size(); ++i) {
        const int jq = q.find_sector(a.sector[i]), jr = r.find_sector(a.sector[i]);
        if(jq < 0 || jr < 0) throw Error(XH_EINVAL, "qr adjoint: Q or R lacks a sector of A");
        const std::size_t m = a.rows[i], n = a.cols[i], k = std::min(m, n);
        if(q.rows[jq] != m || q.cols[jq] != k || r.rows[jr] != k || r.cols[jr] != n)
            throw Error(XH_EINVAL, "qr adjoint: every sector needs the middle dimension min(rows, cols)");
        items.push_back(xh_qr_bwd_item{static_cast<int64_t>(a.offsets[i]), static_cast<int64_t>(q.offsets[jq]), static_cast<int64_t>(r.offsets[jr]),
                                       static_cast<int32_t>(m), static_cast<int32_t>(n)});
    }
    if(items.empty()) return;
    dcheck(A.rt->ctx, xh_qr_bwd(A.rt->ctx, items.data(), static_cast<int64_t>(items.size()), A.data(), Q.data(), R.data(), dQ.data(), dR.data(), dA.data()),
           "qr adjoint");
}

// cutoff_matrices reverse pass (ADTensor.hpp:357-370): the truncated adjoints are added into the
// leading columns of dU, the leading corner of dS and the leading rows of dVdag, sector by sector.
void cutoff_bwd(const DTensor& dUt, const DTensor& dSt, const DTensor& dVt, DTensor& dU, DTensor& dS, DTensor& dVdag)
{
    dU.rt->require_device();
    const TensorStructure &ut = *dUt.st, &st = *dSt.st, &vt = *dVt.st, &u = *dU.st, &s = *dS.st, &v = *dVdag.st;
    std::vector<xh_copy_rect> ru, rs, rv;
    for(std::size_t i = 0; i < ut.sector.size(); ++i) {
        const Q& q = ut.sector[i];
        const int js = st.find_sector(q), jv = vt.find_sector(q);
        const int fu = u.find_sector(q), fs = s.find_sector(q), fv = v.find_sector(q);
        if(js < 0 || jv < 0 || fu < 0 || fs < 0 || fv < 0) throw Error(XH_EINVAL, "cutoff adjoint: sector missing in the full factors");
        const std::size_t kt = ut.cols[i];
        if(ut.rows[i] != u.rows[fu] || kt > u.cols[fu] || st.rows[js] != kt || st.cols[js] != kt || s.rows[fs] < kt || s.cols[fs] < kt ||
           vt.rows[jv] != kt || vt.cols[jv] != v.cols[fv] || kt > v.rows[fv])
            throw Error(XH_EINVAL, "cutoff adjoint: block shapes do not match");
        if(kt == 0) continue;
        ru.push_back(xh_copy_rect{static_cast<int64_t>(ut.offsets[i]), static_cast<int64_t>(ut.rows[i]), static_cast<int64_t>(u.offsets[fu]),
                                  static_cast<int64_t>(u.rows[fu]), static_cast<int32_t>(ut.rows[i]), static_cast<int32_t>(kt), 1.0, 0});
        rs.push_back(xh_copy_rect{static_cast<int64_t>(st.offsets[js]), static_cast<int64_t>(kt), static_cast<int64_t>(s.offsets[fs]),
                                  static_cast<int64_t>(s.rows[fs]), static_cast<int32_t>(kt), static_cast<int32_t>(kt), 1.0, 0});
        rv.push_back(xh_copy_rect{static_cast<int64_t>(vt.offsets[jv]), static_cast<int64_t>(kt), static_cast<int64_t>(v.offsets[fv]),
                                  static_cast<int64_t>(v.rows[fv]), static_cast<int32_t>(kt), static_cast<int32_t>(vt.cols[jv]), 1.0, 0});
    }
    xh_ctx* ctx = dU.rt->ctx;
    if(!ru.empty()) {
        dcheck(ctx, xh_add_rects(ctx, dUt.data(), dU.data(), ru.data(), static_cast<int64_t>(ru.size())), "cutoff adjoint (U)");
        dcheck(ctx, xh_add_rects(ctx, dSt.data(), dS.data(), rs.data(), static_cast<int64_t>(rs.size())), "cutoff adjoint (S)");
        dcheck(ctx, xh_add_rects(ctx, dVt.data(), dVdag.data(), rv.data(), static_cast<int64_t>(rv.size())), "cutoff adjoint (Vdag)");
    }
}

// diag_sqrt / diag_inv reverse passes (ADTensor.hpp:560-590): y is the forward result.
void diag_bwd(const DTensor& y, const DTensor& dy, DTensor& dx, int mode)
{
    y.rt->require_device();
    require_same_blocks(*y.st, *dy.st, "diag adjoint (output)");
    require_same_blocks(*y.st, *dx.st, "diag adjoint (input)");
    std::vector<int64_t> off;
    std::vector<int32_t> m;
    for(std::size_t i = 0; i < y.st->sector.size(); ++i) {
        if(y.st->rows[i] != y.st->cols[i]) throw Error(XH_EINVAL, "diag adjoint needs square blocks");
        off.push_back(static_cast<int64_t>(y.st->offsets[i]));
        m.push_back(static_cast<int32_t>(y.st->rows[i]));
    }
    if(!off.empty())
        dcheck(y.rt->ctx, xh_block_diag_bwd(y.rt->ctx, off.data(), m.data(), static_cast<int64_t>(off.size()), y.data(), dy.data(), dx.data(), mode),
               "diag adjoint");
}

// trace reverse pass (ADTensor.hpp:533-545): dx += g * WeightedIdentity, identity blocks scaled by
// the degeneracy of their sector (Tensor.hpp:238-246).
void trace_bwd(DTensor& dx, double g)
{
    dx.rt->require_device();
    const TensorStructure& s = *dx.st;
    std::vector<int64_t> off;
    std::vector<int32_t> r, c;
    std::vector<double> w;
    for(std::size_t i = 0; i < s.sector.size(); ++i) {
        off.push_back(static_cast<int64_t>(s.offsets[i]));
        r.push_back(static_cast<int32_t>(s.rows[i]));
        c.push_back(static_cast<int32_t>(s.cols[i]));
        w.push_back(g * static_cast<double>(s.sym.degeneracy(s.sector[i])));
    }
    if(!off.empty())
        dcheck(dx.rt->ctx, xh_block_add_identity(dx.rt->ctx, off.data(), r.data(), c.data(), w.data(), static_cast<int64_t>(off.size()), dx.data()),
               "trace adjoint");
}

// maxNorm reverse pass (ADTensor.hpp:450-476): the entry of largest modulus (first in storage
// order) receives g * sign(x).
void max_norm_bwd(const DTensor& x, DTensor& dx, double g)
{
    x.rt->require_device();
    require_same_blocks(*x.st, *dx.st, "maxNorm adjoint");
    if(x.size() == 0) return;
    double v = 0.0;
    int64_t idx = -1;
    dcheck(x.rt->ctx, xh_max_abs_index(x.rt->ctx, x.data(), x.size(), &v, &idx), "maxNorm location");
    if(idx < 0 || static_cast<std::size_t>(idx) >= x.size()) throw Error(XH_EINVAL, "maxNorm adjoint: no maximum");
    double xv = 0.0;
    dcheck(x.rt->ctx, xh_download(x.rt->ctx, &xv, x.data() + idx, sizeof(double)), "maxNorm value");
    dcheck(x.rt->ctx, xh_add_at(x.rt->ctx, dx.data(), idx, std::signbit(xv) ? -g : g), "maxNorm adjoint");
}

}  // namespace xa
