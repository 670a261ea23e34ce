// md_pme.cpp -- FP64 Ewald reciprocal space for the oracle: smooth particle-mesh
// Ewald (Essmann, Perera, Berkowitz, Darden, Lee, Pedersen, J. Chem. Phys. 103,
// 8577 (1995)), the direct Ewald sum over reciprocal vectors that it
// approximates, the correction for excluded pairs and the self term.
//
// TEST INFRASTRUCTURE ONLY (see md_oracle.h).
//
// The reference vendors GROMACS' PME (src/pme.c:829-1340: spread_q_bsplines,
// solve_pme, gather_f_bsplines, make_bsplines, make_bspline_moduli) without
// building it (it needs FFTW and the GROMACS headers), and its lkspace 2 path
// computes only the damped real-space term (src/forces.cu:97-110,
// src/simulation.cu:90-94).  This file restates the method from the paper:
//   Q(k) = sum_i q_i prod_d M_p(u_di - k_d)           (eq. 4.6; u = K (x/L + 1/2))
//   E    = 1/(2 pi V) sum_{m != 0} exp(-pi^2 m^2/b^2)/m^2 B(m) |F(Q)(m)|^2   (4.7)
//   B(m) = prod_d |b_d(m_d)|^2,  1/|b_d(m)|^2 = |sum_{k=0}^{p-2} M_p(k+1) e^{2 pi i m k/K}|^2   (4.4)
//   F_i  = -q_i sum_k (theta_rec * Q)(k) grad_i prod_d M_p(u_di - k_d)          (4.9)
// with m the reciprocal vector (m_d / L_d) of the orthorhombic box, and the
// trace of the reciprocal virial sum_m E_m (1 - 2 pi^2 m^2 / b^2) (eq. 2.6 of
// the paper, traced).  Energies are in kJ/mol with the factor qfac.
//
// The 3D transform is a mixed-radix Cooley-Tukey FFT written here (any length;
// a direct butterfly per prime factor), checked against
// numpy.fft by the tests.
#include "md_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;

// ---------------- FFT ----------------
struct Fft1 {
    int n = 0;
    std::vector<int> fac;
    std::vector<cd> w;   // w[k] = exp(sign 2 pi i k / n)
    Fft1(int n_, int sign) : n(n_) {
        int m = n;
        for (int p = 2; m > 1; p += (p == 2 ? 1 : 2))
            while (m % p == 0) {
                fac.push_back(p);
                m /= p;
            }
        w.resize(n);
        for (int k = 0; k < n; k++) w[k] = std::polar(1.0, sign * 2.0 * kPi * k / n);
    }
    // out[0..len) = DFT of in[0], in[s], ... (len elements); fac from index f
    void rec(const cd* in, int s, cd* out, int len, int f) const {
        if (len == 1) {
            out[0] = in[0];
            return;
        }
        const int p = fac[f], m = len / p;
        for (int q = 0; q < p; q++) rec(in + (size_t)q * s, s * p, out + (size_t)q * m, m, f + 1);
        const int step = n / len;   // w index of exp(sign 2 pi i / len)
        std::vector<cd> t(p);
        for (int k = 0; k < m; k++) {
            for (int q = 0; q < p; q++) t[q] = out[(size_t)q * m + k] * w[((size_t)q * k * step) % n];
            for (int j = 0; j < p; j++) {
                cd acc = 0.0;
                for (int q = 0; q < p; q++) acc += t[q] * w[((size_t)q * j * m * step) % n];
                out[(size_t)j * m + k] = acc;
            }
        }
    }
    void run(cd* a, std::vector<cd>& tmp) const {
        tmp.assign(a, a + n);
        rec(tmp.data(), 1, a, n, 0);
    }
};

// In-place unnormalized 3D DFT of a K0 x K1 x K2 array (last index fastest).
void fft3(cd* a, const int* K, int sign) {
    for (int d = 0; d < 3; d++)
        if (K[d] < 1) return;
    for (int axis = 2; axis >= 0; axis--) {
        const int n = K[axis];
        if (n == 1) continue;
        const size_t stride = axis == 2 ? 1 : (axis == 1 ? (size_t)K[2] : (size_t)K[1] * K[2]);
        const size_t total = (size_t)K[0] * K[1] * K[2];
        const size_t lines = total / n;
        Fft1 plan(n, sign);
#pragma omp parallel
        {
            std::vector<cd> line(n), tmp;
#pragma omp for schedule(static)
            for (long long l = 0; l < (long long)lines; l++) {
                // line l: the start index with the axis coordinate 0
                size_t base;
                if (axis == 2) base = (size_t)l * n;
                else if (axis == 1) base = ((size_t)l / K[2]) * (size_t)K[1] * K[2] + (size_t)l % K[2];
                else base = (size_t)l;
                for (int k = 0; k < n; k++) line[k] = a[base + k * stride];
                plan.run(line.data(), tmp);
                for (int k = 0; k < n; k++) a[base + k * stride] = line[k];
            }
        }
    }
}

// ---------------- B-splines ----------------
// th[j] = M_p(w + j), dth[j] = M_p'(w + j), j = 0..p-1, for 0 <= w < 1:
// M_2(x) = 1 - |x - 1|, M_{k+1}(x) = [x M_k(x) + (k + 1 - x) M_k(x - 1)] / k,
// M_k'(x) = M_{k-1}(x) - M_{k-1}(x - 1)  (Essmann et al. eqs. 4.1-4.2).
void bspline(int p, double w, double* th, double* dth) {
    double a[ORC_PME_MAX_ORDER + 1] = {0};
    a[0] = w;
    a[1] = 1.0 - w;
    for (int k = 2; k < p; k++) {
        if (k == p - 1 && dth)
            for (int j = 0; j < p; j++) dth[j] = (j < k ? a[j] : 0.0) - (j > 0 ? a[j - 1] : 0.0);
        double b[ORC_PME_MAX_ORDER + 1];
        for (int j = 0; j <= k; j++)
            b[j] = ((w + j) * (j < k ? a[j] : 0.0) + (k + 1 - w - j) * (j > 0 ? a[j - 1] : 0.0)) / k;
        for (int j = 0; j <= k; j++) a[j] = b[j];
    }
    if (p == 2 && dth) {
        dth[0] = 1.0;
        dth[1] = -1.0;
    }
    for (int j = 0; j < p; j++) th[j] = a[j];
}

// |b(m)|^2 for m = 0..K-1 (Essmann et al. eq. 4.4)
std::vector<double> bmod2(int p, int K) {
    double th[ORC_PME_MAX_ORDER];
    bspline(p, 0.0, th, nullptr);   // th[j] = M_p(j); M_p(k + 1) = th[k + 1]
    std::vector<double> b(K);
    for (int m = 0; m < K; m++) {
        cd s = 0.0;
        for (int k = 0; k <= p - 2; k++) s += th[k + 1] * std::polar(1.0, 2.0 * kPi * m * k / K);
        b[m] = 1.0 / std::norm(s);
    }
    return b;
}

struct Stencil {
    int k[3][ORC_PME_MAX_ORDER];
    double th[3][ORC_PME_MAX_ORDER], dth[3][ORC_PME_MAX_ORDER];
};

void stencil(const double* x, const double* L, const int* K, int p, Stencil& s) {
    for (int d = 0; d < 3; d++) {
        const double u = K[d] * (x[d] / L[d] + 0.5);
        const double fl = std::floor(u);
        bspline(p, u - fl, s.th[d], s.dth[d]);
        long long k0 = (long long)fl;
        for (int j = 0; j < p; j++) {
            long long k = (k0 - j) % K[d];
            if (k < 0) k += K[d];
            s.k[d][j] = (int)k;
        }
    }
}

inline double mimg(double d, double L) { return d - L * std::rint(d / L); }

}  // namespace

extern "C" void orc_fft3(double* data, const int* K, int sign) {
    fft3(reinterpret_cast<cd*>(data), K, sign < 0 ? -1 : 1);
}

extern "C" void orc_bspline(int order, double w, double* th, double* dth) { bspline(order, w, th, dth); }

extern "C" double orc_pme_recip(const orc_pme* pm, int n, const double* x, const double* q, const float* box3,
                                double qfac, double* F, double* W) {
    const int p = pm->order;
    const int K[3] = {pm->K[0], pm->K[1], pm->K[2]};
    const double L[3] = {box3[0], box3[1], box3[2]};
    const double V = L[0] * L[1] * L[2], b2 = pm->beta * pm->beta;
    const size_t KK = (size_t)K[0] * K[1] * K[2];
    std::vector<cd> Q(KK, 0.0);
    std::vector<Stencil> st(n);
    for (int i = 0; i < n; i++) {
        stencil(x + 3 * (size_t)i, L, K, p, st[i]);
        const Stencil& s = st[i];
        for (int a = 0; a < p; a++)
            for (int b = 0; b < p; b++)
                for (int c = 0; c < p; c++)
                    Q[((size_t)s.k[0][a] * K[1] + s.k[1][b]) * K[2] + s.k[2][c]] +=
                        q[i] * s.th[0][a] * s.th[1][b] * s.th[2][c];
    }
    fft3(Q.data(), K, -1);
    std::vector<double> B[3];
    for (int d = 0; d < 3; d++) B[d] = bmod2(p, K[d]);
    double E = 0.0, Wt = 0.0;
    for (int i0 = 0; i0 < K[0]; i0++) {
        const double m0 = (i0 <= K[0] / 2 ? i0 : i0 - K[0]) / L[0];
        for (int i1 = 0; i1 < K[1]; i1++) {
            const double m1 = (i1 <= K[1] / 2 ? i1 : i1 - K[1]) / L[1];
            for (int i2 = 0; i2 < K[2]; i2++) {
                const double m2 = (i2 <= K[2] / 2 ? i2 : i2 - K[2]) / L[2];
                const size_t idx = ((size_t)i0 * K[1] + i1) * K[2] + i2;
                const double mm = m0 * m0 + m1 * m1 + m2 * m2;
                if (i0 == 0 && i1 == 0 && i2 == 0) {
                    Q[idx] = 0.0;
                    continue;
                }
                const double C = qfac * std::exp(-kPi * kPi * mm / b2) / (kPi * V * mm) * B[0][i0] * B[1][i1] * B[2][i2];
                const double e = 0.5 * C * std::norm(Q[idx]);
                E += e;
                Wt += e * (1.0 - 2.0 * kPi * kPi * mm / b2);
                Q[idx] *= C;
            }
        }
    }
    fft3(Q.data(), K, +1);   // phi(k) = dE/dQ(k)
    if (F) {
        for (int i = 0; i < n; i++) {
            const Stencil& s = st[i];
            double g[3] = {0, 0, 0};
            for (int a = 0; a < p; a++)
                for (int b = 0; b < p; b++)
                    for (int c = 0; c < p; c++) {
                        const double ph = Q[((size_t)s.k[0][a] * K[1] + s.k[1][b]) * K[2] + s.k[2][c]].real();
                        g[0] += ph * s.dth[0][a] * s.th[1][b] * s.th[2][c];
                        g[1] += ph * s.th[0][a] * s.dth[1][b] * s.th[2][c];
                        g[2] += ph * s.th[0][a] * s.th[1][b] * s.dth[2][c];
                    }
            for (int d = 0; d < 3; d++) F[3 * (size_t)i + d] -= q[i] * g[d] * K[d] / L[d];
        }
    }
    if (W) *W = Wt;
    return E;
}

extern "C" double orc_ewald_recip(int n, const double* x, const double* q, const float* box3, double beta, double qfac,
                                  int kmax, double* F, double* W) {
    const double L[3] = {box3[0], box3[1], box3[2]};
    const double V = L[0] * L[1] * L[2], b2 = beta * beta;
    double E = 0.0, Wt = 0.0;
    std::vector<double> f(3 * (size_t)n, 0.0);
    std::vector<cd> e(n);
    for (int a = -kmax; a <= kmax; a++)
        for (int b = -kmax; b <= kmax; b++)
            for (int c = -kmax; c <= kmax; c++) {
                if (a == 0 && b == 0 && c == 0) continue;
                const double m[3] = {a / L[0], b / L[1], c / L[2]};
                const double mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                const double g = qfac * std::exp(-kPi * kPi * mm / b2) / mm;
                cd S = 0.0;
                for (int i = 0; i < n; i++) {
                    const double ph = 2.0 * kPi * (m[0] * x[3 * i] + m[1] * x[3 * i + 1] + m[2] * x[3 * i + 2]);
                    e[i] = std::polar(1.0, ph);
                    S += q[i] * e[i];
                }
                const double em = g * std::norm(S) / (2.0 * kPi * V);
                E += em;
                Wt += em * (1.0 - 2.0 * kPi * kPi * mm / b2);
                // F_i = -dE/dx_i = (2 q_i / V) g m Im(conj(S) e_i)
                for (int i = 0; i < n; i++) {
                    const double im = (std::conj(S) * e[i]).imag();
                    for (int d = 0; d < 3; d++) f[3 * (size_t)i + d] += 2.0 * q[i] / V * g * m[d] * im;
                }
            }
    if (F)
        for (size_t k = 0; k < f.size(); k++) F[k] += f[k];
    if (W) *W = Wt;
    return E;
}

extern "C" double orc_pme_excl(int n, const double* x, const double* q, const float* box3, const int* excl, double beta,
                               double qfac, double* F, double* W) {
    const double L[3] = {box3[0], box3[1], box3[2]};
    const double c2 = 2.0 * beta / std::sqrt(kPi);
    double E = 0.0, Wt = 0.0;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) {
            const int j = excl[4 * (size_t)i + k];
            if (j < 0) continue;
            double d[3];
            for (int t = 0; t < 3; t++) d[t] = mimg(x[3 * (size_t)i + t] - x[3 * (size_t)j + t], L[t]);
            const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const double qq = qfac * q[i] * q[j];
            const double erfr = std::erf(beta * r);
            // U = -qq erf(b r)/r; F_i = -dU/dr d/r = qq [c2 e^{-b^2 r^2}/r - erf(b r)/r^2] d/r
            const double fr = qq * (c2 * std::exp(-beta * beta * r * r) / r - erfr / (r * r)) / r;
            if (F)
                for (int t = 0; t < 3; t++) F[3 * (size_t)i + t] += fr * d[t];
            E -= 0.5 * qq * erfr / r;   // each pair is listed from both atoms
            Wt += 0.5 * fr * r * r;
        }
    if (W) *W = Wt;
    return E;
}

extern "C" double orc_ewald_self(int n, const double* q, const float* box3, double beta, double qfac, double* W) {
    double s2 = 0.0, s1 = 0.0;
    for (int i = 0; i < n; i++) {
        s2 += q[i] * q[i];
        s1 += q[i];
    }
    const double V = (double)box3[0] * box3[1] * box3[2];
    const double ebg = -qfac * kPi * s1 * s1 / (2.0 * V * beta * beta);   // neutralizing background
    if (W) *W = 3.0 * ebg;
    return -qfac * beta / std::sqrt(kPi) * s2 + ebg;
}
