// Runtime symmetry descriptor for the block-sparse tensor core.
//
// The reference encodes symmetries as compile-time structs (include/Xped/Symmetry/U1.hpp,
// SU2.hpp, S1xS2.hpp, CombSym.hpp).  Here a symmetry is a runtime value holding up to two
// components, each U(1) (bosonic/spin or fermionic charge) or SU(2) (labels q = 2S+1).
// Every coefficient below is the product of the per-component coefficients, exactly as
// S1xS2.hpp:105-308 composes them.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

namespace xa {

constexpr int MAXQ = 2;

enum SymKind : int {
    KIND_U1 = 1,   // U1<SpinU1>/U1<ChargeU1>: abelian, bosonic        (U1.hpp:40-226)
    KIND_FU1 = 2,  // U1<FChargeU1>: abelian, fermionic parity = q odd  (U1.hpp:130-134,174-182)
    KIND_SU2 = 3,  // SU2<SpinSU2>: non-abelian, q = 2S+1               (SU2.hpp:108-177)
};

struct Q {
    std::array<int, MAXQ> v{0, 0};
    bool operator==(const Q& o) const { return v == o.v; }
    bool operator!=(const Q& o) const { return v != o.v; }
    // qarray operator< is the lexicographic std::array comparison (qarray.hpp:111-114)
    bool operator<(const Q& o) const { return v < o.v; }
    bool operator>(const Q& o) const { return v > o.v; }
};

// Wigner symbols from the Racah formulas (integer arguments are 2j, i.e. q-1).
double wigner_3j(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3);
double wigner_6j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6);
double wigner_9j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6, int tj7, int tj8, int tj9);

inline double phase(int x) { return (x % 2) ? -1.0 : 1.0; }  // functions.hpp:79-84

struct Sym {
    int nq = 0;
    std::array<int, MAXQ> kind{0, 0};

    static Sym make(int k0, int k1 = 0)
    {
        Sym s;
        s.kind[0] = k0;
        s.kind[1] = k1;
        s.nq = (k0 ? 1 : 0) + (k1 ? 1 : 0);
        return s;
    }

    bool operator==(const Sym& o) const { return nq == o.nq && kind == o.kind; }

    bool any_non_abelian() const
    {
        for(int c = 0; c < nq; ++c)
            if(kind[c] == KIND_SU2) return true;
        return false;
    }
    bool any_fermionic() const
    {
        for(int c = 0; c < nq; ++c)
            if(kind[c] == KIND_FU1) return true;
        return false;
    }

    Q vacuum() const
    {
        Q q;
        for(int c = 0; c < nq; ++c) q.v[c] = (kind[c] == KIND_SU2) ? 1 : 0;
        return q;
    }
    Q conj(const Q& q) const
    {
        Q out = q;
        for(int c = 0; c < nq; ++c)
            if(kind[c] != KIND_SU2) out.v[c] = -q.v[c];
        return out;
    }
    int degeneracy(const Q& q) const
    {
        int d = 1;
        for(int c = 0; c < nq; ++c)
            if(kind[c] == KIND_SU2) d *= q.v[c];
        return d;
    }
    // SymBase::compare on a single qType is operator< (SymBase.hpp:94)
    bool compare(const Q& a, const Q& b) const { return a < b; }

    // basis_combine: U1.hpp:189-194, SU2.cpp:18-27, S1xS2.hpp:228-240 (component 0 outer loop)
    std::vector<Q> combine(const Q& a, const Q& b) const
    {
        std::vector<Q> out;
        if(nq == 0) {
            out.push_back(Q{});
            return out;
        }
        std::vector<int> c0 = combine_comp(0, a.v[0], b.v[0]);
        if(nq == 1) {
            for(int x : c0) {
                Q q;
                q.v[0] = x;
                out.push_back(q);
            }
            return out;
        }
        std::vector<int> c1 = combine_comp(1, a.v[1], b.v[1]);
        for(int x : c0)
            for(int y : c1) {
                Q q;
                q.v[0] = x;
                q.v[1] = y;
                out.push_back(q);
            }
        return out;
    }

    bool triangle(const Q& a, const Q& b, const Q& c) const
    {
        for(int k = 0; k < nq; ++k)
            if(!triangle_comp(k, a.v[k], b.v[k], c.v[k])) return false;
        return true;
    }

    double coeff_swap(const Q& ql, const Q& qr, const Q& qf) const
    {
        double out = 1.0;
        for(int c = 0; c < nq; ++c) out *= swap_comp(c, ql.v[c], qr.v[c], qf.v[c]);
        return out;
    }
    double coeff_recouple(const Q& q1, const Q& q2, const Q& q3, const Q& Qt, const Q& Q12, const Q& Q23) const
    {
        double out = 1.0;
        for(int c = 0; c < nq; ++c) out *= recouple_comp(c, q1.v[c], q2.v[c], q3.v[c], Qt.v[c], Q12.v[c], Q23.v[c]);
        return out;
    }
    double coeff_turn(const Q& ql, const Q& qr, const Q& qf) const
    {
        double out = 1.0;
        for(int c = 0; c < nq; ++c) out *= turn_comp(c, ql.v[c], qr.v[c], qf.v[c]);
        return out;
    }
    double coeff_FS(const Q& q) const
    {
        double out = 1.0;
        for(int c = 0; c < nq; ++c)
            if(kind[c] == KIND_SU2) out *= (q.v[c] % 2 == 0) ? -1.0 : 1.0;  // SU2.cpp:35-39
        return out;
    }
    double coeff_twist(const Q& q) const
    {
        double out = 1.0;
        for(int c = 0; c < nq; ++c)
            if(kind[c] == KIND_FU1) out *= (q.v[c] % 2 != 0) ? -1.0 : 1.0;  // U1.hpp:130-134
        return out;
    }

private:
    std::vector<int> combine_comp(int c, int a, int b) const
    {
        if(kind[c] == KIND_SU2) {
            std::vector<int> out;
            int qmin = std::abs(a - b) + 1, qmax = std::abs(a + b) - 1;
            for(int i = qmin; i <= qmax; i += 2) out.push_back(i);
            return out;
        }
        return {a + b};
    }
    bool triangle_comp(int c, int a, int b, int q3) const
    {
        if(kind[c] == KIND_SU2) return (q3 - 1 >= std::abs(a - b)) && (q3 - 1 <= a + b - 2);  // SU2.cpp:114-119
        return a + b == q3;                                                                   // U1.hpp:221-226
    }
    double swap_comp(int c, int ql, int qr, int qf) const
    {
        if(kind[c] == KIND_SU2) {  // SU2.hpp:169-177 (bosonic SU2)
            return triangle_comp(c, ql, qr, qf) ? phase((ql + qr - qf - 1) / 2) : 0.0;
        }
        double sign = 1.0;
        if(kind[c] == KIND_FU1) sign = ((ql % 2 != 0) && (qr % 2 != 0)) ? -1.0 : 1.0;  // U1.hpp:174-182
        return triangle_comp(c, ql, qr, qf) ? sign : 0.0;
    }
    double sixj_comp(int c, int q1, int q2, int q3, int q4, int q5, int q6) const
    {
        if(kind[c] == KIND_SU2) return wigner_6j(q1 - 1, q2 - 1, q3 - 1, q4 - 1, q5 - 1, q6 - 1);
        // U1.hpp:214-219
        if(triangle_comp(c, q1, q2, q3) && triangle_comp(c, q1, q6, q5) && triangle_comp(c, q2, q4, q6) &&
           triangle_comp(c, q3, q4, q5))
            return 1.0;
        return 0.0;
    }
    double recouple_comp(int c, int q1, int q2, int q3, int Qt, int Q12, int Q23) const
    {
        if(kind[c] == KIND_SU2) {  // SU2.hpp:164-167; phase() takes an int, the argument sum is even
            double s = std::sqrt(static_cast<double>(Q12 * Q23));
            return s * phase(static_cast<int>((q1 + q2 + q3 + Qt) / 2.)) * sixj_comp(c, q1, q2, Q12, q3, Qt, Q23);
        }
        return sixj_comp(c, q1, q2, Q12, q3, Qt, Q23);  // U1.hpp:163-166
    }
    double turn_comp(int c, int ql, int qr, int qf) const
    {
        if(kind[c] == KIND_SU2) {  // SU2.hpp:143-146
            return triangle_comp(c, ql, qr, qf)
                       ? swap_comp(c, ql, qr, qf) * std::sqrt(static_cast<double>(qf) / static_cast<double>(ql))
                       : 0.0;
        }
        return triangle_comp(c, ql, qr, qf) ? 1.0 : 0.0;  // U1.hpp:159
    }
};

}  // namespace xa
