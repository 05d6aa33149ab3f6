// Wigner 3j/6j/9j symbols for SU(2) recoupling.
//
// The reference obtains them from GSL (include/Xped/Symmetry/SU2Wrappers.hpp:66-82, arguments
// are 2j = q-1).  Here they are evaluated from the Racah single-sum formulas (Edmonds,
// "Angular Momentum in Quantum Mechanics", eqs. 3.7.3 and 6.3.7; 9j as the 6j-product sum,
// eq. 6.4.3) in extended precision.  Factorials up to 170! are exact-range long doubles and the
// alternating sums are accumulated in long double, which keeps the result within a few ulps of
// the exact value for the spins used by the tensor network code (2j <= 40).
#include <cmath>
#include <mutex>

#include "sym.hpp"

namespace xa {

namespace {
constexpr int NFACT = 171;
long double g_fact[NFACT];
std::once_flag g_fact_once;

void init_fact()
{
    g_fact[0] = 1.0L;
    for(int i = 1; i < NFACT; ++i) g_fact[i] = g_fact[i - 1] * static_cast<long double>(i);
}

inline long double fact(int n) { return g_fact[n]; }

// Triangle coefficient Delta(abc) for doubled arguments; returns -1 when the triad is invalid.
long double delta2(int ta, int tb, int tc)
{
    int s1 = ta + tb - tc, s2 = ta - tb + tc, s3 = -ta + tb + tc, s4 = ta + tb + tc;
    if(s1 < 0 || s2 < 0 || s3 < 0 || (s4 & 1)) return -1.0L;
    return fact(s1 / 2) * fact(s2 / 2) * fact(s3 / 2) / fact(s4 / 2 + 1);
}
}  // namespace

double wigner_3j(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3)
{
    std::call_once(g_fact_once, init_fact);
    if(tj1 < 0 || tj2 < 0 || tj3 < 0) return 0.0;
    if(tm1 + tm2 + tm3 != 0) return 0.0;
    if(std::abs(tm1) > tj1 || std::abs(tm2) > tj2 || std::abs(tm3) > tj3) return 0.0;
    if(((tj1 + tm1) & 1) || ((tj2 + tm2) & 1) || ((tj3 + tm3) & 1)) return 0.0;
    long double d = delta2(tj1, tj2, tj3);
    if(d < 0) return 0.0;
    // all half-differences below are integers once the parities above hold
    int j1pm1 = (tj1 + tm1) / 2, j1mm1 = (tj1 - tm1) / 2;
    int j2pm2 = (tj2 + tm2) / 2, j2mm2 = (tj2 - tm2) / 2;
    int j3pm3 = (tj3 + tm3) / 2, j3mm3 = (tj3 - tm3) / 2;
    int a1 = (tj3 - tj2 + tm1) / 2;  // j3 - j2 + m1
    int a2 = (tj3 - tj1 - tm2) / 2;  // j3 - j1 - m2
    int a3 = (tj1 + tj2 - tj3) / 2;  // j1 + j2 - j3
    int a4 = j1mm1;                  // j1 - m1
    int a5 = j2pm2;                  // j2 + m2
    int kmin = 0;
    if(-a1 > kmin) kmin = -a1;
    if(-a2 > kmin) kmin = -a2;
    int kmax = a3;
    if(a4 < kmax) kmax = a4;
    if(a5 < kmax) kmax = a5;
    long double sum = 0.0L;
    for(int k = kmin; k <= kmax; ++k) {
        long double term = 1.0L / (fact(k) * fact(a1 + k) * fact(a2 + k) * fact(a3 - k) * fact(a4 - k) * fact(a5 - k));
        sum += (k & 1) ? -term : term;
    }
    long double pre = std::sqrt(d * fact(j1pm1) * fact(j1mm1) * fact(j2pm2) * fact(j2mm2) * fact(j3pm3) * fact(j3mm3));
    int ph = (tj1 - tj2 - tm3) / 2;  // (-1)^(j1-j2-m3)
    long double out = pre * sum;
    if(ph % 2) out = -out;
    return static_cast<double>(out);
}

double wigner_6j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6)
{
    std::call_once(g_fact_once, init_fact);
    if(tj1 < 0 || tj2 < 0 || tj3 < 0 || tj4 < 0 || tj5 < 0 || tj6 < 0) return 0.0;
    long double d1 = delta2(tj1, tj2, tj3), d2 = delta2(tj1, tj5, tj6), d3 = delta2(tj4, tj2, tj6), d4 = delta2(tj4, tj5, tj3);
    if(d1 < 0 || d2 < 0 || d3 < 0 || d4 < 0) return 0.0;
    int a1 = (tj1 + tj2 + tj3) / 2, a2 = (tj1 + tj5 + tj6) / 2, a3 = (tj4 + tj2 + tj6) / 2, a4 = (tj4 + tj5 + tj3) / 2;
    int b1 = (tj1 + tj2 + tj4 + tj5) / 2, b2 = (tj2 + tj3 + tj5 + tj6) / 2, b3 = (tj3 + tj1 + tj6 + tj4) / 2;
    int tmin = a1;
    if(a2 > tmin) tmin = a2;
    if(a3 > tmin) tmin = a3;
    if(a4 > tmin) tmin = a4;
    int tmax = b1;
    if(b2 < tmax) tmax = b2;
    if(b3 < tmax) tmax = b3;
    long double sum = 0.0L;
    for(int t = tmin; t <= tmax; ++t) {
        long double term = fact(t + 1) / (fact(t - a1) * fact(t - a2) * fact(t - a3) * fact(t - a4) * fact(b1 - t) * fact(b2 - t) * fact(b3 - t));
        sum += (t & 1) ? -term : term;
    }
    return static_cast<double>(std::sqrt(d1 * d2 * d3 * d4) * sum);
}

double wigner_9j(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6, int tj7, int tj8, int tj9)
{
    // {j1 j2 j3; j4 j5 j6; j7 j8 j9} = sum_x (-1)^{2x} (2x+1) {j1 j4 j7; j8 j9 x}{j2 j5 j8; j4 x j6}{j3 j6 j9; x j1 j2}
    int xmin = std::abs(tj1 - tj9);
    int t = std::abs(tj4 - tj8);
    if(t > xmin) xmin = t;
    t = std::abs(tj2 - tj6);
    if(t > xmin) xmin = t;
    int xmax = tj1 + tj9;
    if(tj4 + tj8 < xmax) xmax = tj4 + tj8;
    if(tj2 + tj6 < xmax) xmax = tj2 + tj6;
    long double sum = 0.0L;
    for(int tx = xmin; tx <= xmax; tx += 2) {
        long double w = static_cast<long double>(wigner_6j(tj1, tj4, tj7, tj8, tj9, tx)) * wigner_6j(tj2, tj5, tj8, tj4, tx, tj6) *
                        wigner_6j(tj3, tj6, tj9, tx, tj1, tj2);
        long double term = (tx + 1) * w;
        sum += (tx & 1) ? -term : term;
    }
    return static_cast<double>(sum);
}

}  // namespace xa
