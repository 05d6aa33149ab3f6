// Reverse-pass operations of the AD tensor (include/Xped/AD/ADTensor.hpp) on device.
#pragma once

#include "tensor.hpp"

namespace xa {

// operator* reverse pass (ADTensor.hpp:745-746): dA += dC * B^dagger, dB += A^dagger * dC.
// Either output may be null.  dA and dB must have the structures of A and B.
void mul_bwd(const DTensor& A, const DTensor& B, const DTensor& dC, DTensor* dA, DTensor* dB, int precision);
// full_svd_impl reverse pass (ADTensor.hpp:262-320), accumulated into dA (structure of the
// decomposed tensor).  U, S, Vdag are the untruncated factors and dU, dS, dVdag their adjoints.
void svd_bwd(const DTensor& U, const DTensor& S, const DTensor& Vdag, const DTensor& dU, const DTensor& dS, const DTensor& dVdag, DTensor& dA);
// cutoff_matrices reverse pass (ADTensor.hpp:357-370): adds the adjoints of the truncated
// factors into the leading columns / corner / rows of the adjoints of the full factors.
void cutoff_bwd(const DTensor& dUt, const DTensor& dSt, const DTensor& dVt, DTensor& dU, DTensor& dS, DTensor& dVdag);

// tQR reverse pass (no counterpart in the reference, whose AD tensor records no tQR): with
// (Q, R) = tQR(A), every sector needs the middle dimension min(rows, cols) and a full-rank R;
// dA += the QR adjoint of the sector (xh_qr_bwd).
void qr_bwd(const DTensor& A, const DTensor& Q, const DTensor& R, const DTensor& dQ, const DTensor& dR, DTensor& dA);

// diag_sqrt (mode 0) / diag_inv (mode 1) / diag_sqrt().diag_inv() (mode 2) reverse passes
// (ADTensor.hpp:560-590); y is the result.
void diag_bwd(const DTensor& y, const DTensor& dy, DTensor& dx, int mode);
// trace reverse pass (ADTensor.hpp:533-545): dx += g * WeightedIdentity.
void trace_bwd(DTensor& dx, double g);
// maxNorm reverse pass (ADTensor.hpp:450-476).
void max_norm_bwd(const DTensor& x, DTensor& dx, double g);

}  // namespace xa
