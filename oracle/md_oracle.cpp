/*
 * md_oracle.cpp -- FP64 CPU restatement of the md_gpu reference MD step.
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the timed CPU
 * baseline).  Nothing in the product links or loads this file's library.
 *
 * Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).  Contraction is
 * off so that the canonical FP32 predicates below round exactly as written;
 * every FMA in them is an explicit fmaf().
 *
 * Pinning (each function, and the test that pins it; tests/test_oracle_*.py):
 *   orc_canonical_thresholds  formula check against numpy float32 arithmetic
 *   orc_rsq32                 brute-force numpy float32 restatement on random
 *                             pairs incl. |dx| near L/2 (test_oracle_nlist)
 *   orc_nlist_brute/_cells    set equality brute vs cells, symmetry
 *                             j in N(i) <=> i in N(j), self/exclusion absent,
 *                             scipy cKDTree sets on off-boundary pairs
 *   orc_pair_forces           LJ: F = 0 at r = 2^(1/6) sigma, U(sigma) = U(rc) shift,
 *                             -dU/dr by central FD; Wolf: F(rc) = 0, alpha -> 0
 *                             limit, A&S vs exact erfc within 1.5e-7, -dU/dr by FD;
 *                             RF closed form; sum_i F_i = 0; virial vs FD of U
 *                             under uniform scaling
 *   orc_bond_forces           FD of k(r-r0)^2, sum F = 0
 *   orc_angle_forces          FD of k(theta-theta0)^2, sum F = 0, zero torque
 *   orc_md_run                constant-force parabola (LF/VV), free flight,
 *                             barostat: zero compressibility = NVT bit for bit,
 *                             ideal-gas volume N_f kT0 / (3 P0), pressure response,
 *                             harmonic-dimer period, VV time reversibility,
 *                             NVE drift bound, Berendsen relaxation rate,
 *                             Nose-Hoover conserved quantity and <T>
 *   orc_fm_loss_grad          central finite differences of the loss (h = 1e-4
 *                             relative), zero gradient at the target parameters
 *   orc_settle                exact O-H / H-H distances, agreement with orc_shake
 *                             (an independent iterative method), the centre of mass
 *                             kept, identity on positions already satisfying the
 *                             constraints after a rigid motion; in orc_md_run: NVE
 *                             energy conservation of rigid water
 * The FP32 device intrinsics of the reference (rsqrtf, __fdividef, exp2f, acosf
 * on compute capability 1.1, src/pair_forces.cu:175-181, src/bonded_forces.cu:165)
 * have no reference test vectors: parity at that boundary is unpinned and is
 * replaced by FP64 libm here.
 */
#include "md_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// A&S 7.1.26 constants, reference src/simulation.h:39-44.
const double AS_A1 = 0.254829592, AS_A2 = -0.284496736, AS_A3 = 1.421413741;
const double AS_A4 = -1.453152027, AS_A5 = 1.061405429, AS_P = 0.3275911;
const double SQRT_PI = 1.7724538509055160273;
// Temperature conversion of the reference: T = KE2/(3N-3)*1000/8.314 (src/integrator.cu:155)
const double KB_REF = 8.314e-3;  // kJ/mol/K

// ---------------- canonical FP32 geometry ----------------
inline float minimg32(float d, float L, float Linv) {
    float s = rintf(d * Linv);          // round-half-even, src/neighborlist.cu:271-273
    return fmaf(-L, s, d);              // L*s exact for s in {-1,0,1}
}
inline float rsq32(float dx, float dy, float dz) {
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}
inline float shift32(float d, float Linv) { return rintf(d * Linv); }

// ---------------- erfc pieces (FP64) ----------------
// E(r) = erfc(alpha r) via A&S (src/pair_forces.cu:178-181) or libm.
inline double erfc_part(double r, double alpha, int exact) {
    if (exact) return std::erfc(alpha * r);
    double t = 1.0 / (1.0 + AS_P * alpha * r);
    double poly = ((((AS_A5 * t + AS_A4) * t + AS_A3) * t + AS_A2) * t + AS_A1) * t;
    return poly * std::exp(-alpha * alpha * r * r);
}
// A(r) = erfc(alpha r) + (2 alpha/sqrt(pi)) r exp(-alpha^2 r^2)
inline double a_part(double r, double alpha, int exact) {
    return erfc_part(r, alpha, exact) + 2.0 * alpha / SQRT_PI * r * std::exp(-alpha * alpha * r * r);
}

struct CoulConst {
    double wolf;     // A(rc)/rc for Wolf (src/forces.cu:104-109), 0 otherwise
    double ecut;     // E(rc)/rc (energy shift)
    double reacf;    // 2(eps-1)/((2eps+1) rc^3), src/simulation.cu:86
    double crf;      // RF energy constant so that U(rc) = 0
};
CoulConst coul_const(const orc_ff* ff) {
    CoulConst c{0, 0, 0, 0};
    double rc = ff->rc;
    if (ff->coulomb == ORC_COUL_WOLF) {
        c.wolf = a_part(rc, ff->alpha, ff->erfc_exact) / rc;
        c.ecut = erfc_part(rc, ff->alpha, ff->erfc_exact) / rc;
    } else if (ff->coulomb == ORC_COUL_RF) {
        c.reacf = 2.0 * (ff->eps_rf - 1.0) / (2.0 * ff->eps_rf + 1.0) / (rc * rc * rc);
        c.crf = 1.0 / rc + 0.5 * c.reacf * rc * rc;
    }
    return c;
}

// ff_coul / (qfac qi qj) and U_coul / (qfac qi qj) at distance r.
inline void coul_terms(const orc_ff* ff, const CoulConst& cc, double r, double* fr, double* ur) {
    double rsq = r * r;
    switch (ff->coulomb) {
    case ORC_COUL_RF:   // src/pair_forces.cu:88-91
        *fr = (1.0 / r - cc.reacf * rsq) / rsq;
        if (ur) *ur = 1.0 / r + 0.5 * cc.reacf * rsq - cc.crf;
        break;
    case ORC_COUL_WOLF: {  // src/pair_forces.cu:178-183
        double A = a_part(r, ff->alpha, ff->erfc_exact);
        *fr = (A / r - cc.wolf) / rsq;
        if (ur) *ur = erfc_part(r, ff->alpha, ff->erfc_exact) / r - cc.ecut + cc.wolf * std::log(r / ff->rc);
        break;
    }
    case ORC_COUL_DAMPED: {  // lkspace 2: same kernel with wolf = 0 (src/forces.cu:104-110)
        double A = a_part(r, ff->alpha, ff->erfc_exact);
        *fr = (A / r) / rsq;
        if (ur) *ur = erfc_part(r, ff->alpha, ff->erfc_exact) / r;
        break;
    }
    default:
        *fr = 0.0;
        if (ur) *ur = 0.0;
    }
}

inline double lj_force(double l1, double l2, double rsq) {   // src/pair_forces.cu:190-193
    double ri2 = 1.0 / rsq, ri8 = ri2 * ri2 * ri2 * ri2;
    return (l1 * ri2 * ri2 * ri2 - l2) * ri8;
}
inline double lj_energy(double l1, double l2, double rsq) {  // 4eps[(s/r)^12-(s/r)^6] = l1/12 r^-12 - l2/6 r^-6
    double ri6 = 1.0 / (rsq * rsq * rsq);
    return (l1 / 12.0) * ri6 * ri6 - (l2 / 6.0) * ri6;
}

// ---------------- neighbor lists ----------------
void finish_lists(int n, int cap, std::vector<std::vector<int>>& L, int* counts, int* nbrs, int* maxc) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        auto& v = L[i];
        std::sort(v.begin(), v.end());
        counts[i] = (int)v.size();
        m = std::max(m, counts[i]);
        int c = std::min(cap, counts[i]);
        for (int k = 0; k < c; k++) nbrs[(size_t)i * cap + k] = v[k];
    }
    *maxc = m;
}

inline bool excluded(const int* excl, int i, int j) {
    if (!excl) return false;
    const int* e = excl + 4 * (size_t)i;
    return e[0] == j || e[1] == j || e[2] == j || e[3] == j;
}

struct Cells {
    int nc[3];
    std::vector<int> start, list;  // CSR over cells
};
// Bin positions into nc^3 cells of edge L/nc >= rl (returns false if nc < 3 in some dim).
bool build_cells(int n, const double* pos, const double* L, double rl, Cells& c) {
    for (int d = 0; d < 3; d++) {
        c.nc[d] = (int)std::floor(L[d] / rl);
        if (c.nc[d] < 3) return false;
    }
    int ncell = c.nc[0] * c.nc[1] * c.nc[2];
    std::vector<int> cell(n);
    c.start.assign(ncell + 1, 0);
    for (int i = 0; i < n; i++) {
        int cc[3];
        for (int d = 0; d < 3; d++) {
            double x = pos[3 * (size_t)i + d];
            double xw = x - L[d] * std::floor((x + 0.5 * L[d]) / L[d]);  // into [-L/2, L/2)
            int k = (int)std::floor((xw + 0.5 * L[d]) / L[d] * c.nc[d]);
            cc[d] = std::min(std::max(k, 0), c.nc[d] - 1);
        }
        cell[i] = cc[0] + c.nc[0] * (cc[1] + c.nc[1] * cc[2]);
        c.start[cell[i] + 1]++;
    }
    for (int k = 0; k < ncell; k++) c.start[k + 1] += c.start[k];
    c.list.assign(n, 0);
    std::vector<int> fill(c.start.begin(), c.start.end() - 1);
    for (int i = 0; i < n; i++) c.list[fill[cell[i]]++] = i;
    return true;
}
template <class F>
void for_cell_neighbors(const Cells& c, const double* xi, const double* L, F&& f) {
    int cc[3];
    for (int d = 0; d < 3; d++) {
        double xw = xi[d] - L[d] * std::floor((xi[d] + 0.5 * L[d]) / L[d]);
        int k = (int)std::floor((xw + 0.5 * L[d]) / L[d] * c.nc[d]);
        cc[d] = std::min(std::max(k, 0), c.nc[d] - 1);
    }
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int x = (cc[0] + dx + c.nc[0]) % c.nc[0];
                int y = (cc[1] + dy + c.nc[1]) % c.nc[1];
                int z = (cc[2] + dz + c.nc[2]) % c.nc[2];
                int k = x + c.nc[0] * (y + c.nc[1] * z);
                for (int p = c.start[k]; p < c.start[k + 1]; p++) f(c.list[p]);
            }
}

}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_canonical_thresholds(float rc, float skin, const float* box3,
                              float* rc2, float* rl2, float* skin2q, float* linv3) {
    float r2 = rc * rc;
    *rc2 = r2;
    double rl = (double)skin + std::sqrt((double)r2);
    *rl2 = (float)(rl * rl);
    *skin2q = skin * skin / 4.0f;
    for (int d = 0; d < 3; d++) linv3[d] = (float)(1.0 / (double)box3[d]);
}

float orc_rsq32(const float* xi, const float* xj, const float* box3, const float* linv3) {
    float dx = minimg32(xi[0] - xj[0], box3[0], linv3[0]);
    float dy = minimg32(xi[1] - xj[1], box3[1], linv3[1]);
    float dz = minimg32(xi[2] - xj[2], box3[2], linv3[2]);
    return rsq32(dx, dy, dz);
}

int orc_nlist_brute(int n, const float* pos, const float* box3, float rl2,
                    const int* excl, int cap, int* counts, int* nbrs) {
    float linv[3];
    for (int d = 0; d < 3; d++) linv[d] = (float)(1.0 / (double)box3[d]);
    std::vector<std::vector<int>> L(n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            if (j == i || excluded(excl, i, j)) continue;
            if (orc_rsq32(pos + 3 * (size_t)i, pos + 3 * (size_t)j, box3, linv) < rl2) L[i].push_back(j);
        }
    }
    int m;
    finish_lists(n, cap, L, counts, nbrs, &m);
    return m;
}

int orc_nlist_cells(int n, const float* pos, const float* box3, float rl2,
                    const int* excl, int cap, int* counts, int* nbrs) {
    double Ld[3] = {box3[0], box3[1], box3[2]};
    std::vector<double> pd(3 * (size_t)n);
    for (size_t k = 0; k < 3 * (size_t)n; k++) pd[k] = pos[k];
    Cells c;
    // cell edge must cover the FP32 list radius; pad slightly for the float/double mismatch
    if (!build_cells(n, pd.data(), Ld, std::sqrt((double)rl2) * (1.0 + 1e-6), c))
        return orc_nlist_brute(n, pos, box3, rl2, excl, cap, counts, nbrs);
    float linv[3];
    for (int d = 0; d < 3; d++) linv[d] = (float)(1.0 / (double)box3[d]);
    std::vector<std::vector<int>> L(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        for_cell_neighbors(c, pd.data() + 3 * (size_t)i, Ld, [&](int j) {
            if (j == i || excluded(excl, i, j)) return;
            if (orc_rsq32(pos + 3 * (size_t)i, pos + 3 * (size_t)j, box3, linv) < rl2) L[i].push_back(j);
        });
    }
    int m;
    finish_lists(n, cap, L, counts, nbrs, &m);
    return m;
}

void orc_pair_forces(int n, const float* pos, const int* type, const double* charge,
                     const float* box3, const orc_ff* ff, float rc2,
                     int cap, const int* counts, const int* nbrs,
                     double* F, double* energy, double* virial) {
    float linv[3];
    for (int d = 0; d < 3; d++) linv[d] = (float)(1.0 / (double)box3[d]);
    const int nt = ff->ntype;
    CoulConst cc = coul_const(ff);
    const double rc = ff->rc;
    double elj = 0.0, ecoul = 0.0, vir = 0.0;
    const bool want_e = energy != nullptr || virial != nullptr;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : elj, ecoul, vir)
    for (int i = 0; i < n; i++) {
        const float* xi = pos + 3 * (size_t)i;
        double f[3] = {0, 0, 0};
        int cnt = std::min(counts[i], cap);
        for (int k = 0; k < cnt; k++) {
            int j = nbrs[(size_t)i * cap + k];
            const float* xj = pos + 3 * (size_t)j;
            // canonical FP32 geometry decides membership and the periodic image
            float d32[3], s[3];
            for (int d = 0; d < 3; d++) {
                float dd = xi[d] - xj[d];
                s[d] = shift32(dd, linv[d]);
                d32[d] = fmaf(-box3[d], s[d], dd);
            }
            float r2f = rsq32(d32[0], d32[1], d32[2]);
            if (r2f > rc2) continue;            // src/pair_forces.cu:195
            double d64[3];
            for (int d = 0; d < 3; d++) d64[d] = ((double)xi[d] - (double)xj[d]) - (double)box3[d] * (double)s[d];
            double rsq = d64[0] * d64[0] + d64[1] * d64[1] + d64[2] * d64[2];
            double r = std::sqrt(rsq);
            int tp = type[i] * nt + type[j];
            double qq = ff->qfac * charge[i] * charge[j];
            double fr, ur;
            coul_terms(ff, cc, r, &fr, want_e ? &ur : nullptr);
            double ffv = qq * fr + lj_force(ff->lj1[tp], ff->lj2[tp], rsq);
            for (int d = 0; d < 3; d++) f[d] += ffv * d64[d];
            if (want_e) {
                double rc2d = rc * rc;
                elj += 0.5 * (lj_energy(ff->lj1[tp], ff->lj2[tp], rsq) - lj_energy(ff->lj1[tp], ff->lj2[tp], rc2d));
                ecoul += 0.5 * qq * ur;
                vir += 0.5 * ffv * rsq;
            }
        }
        for (int d = 0; d < 3; d++) F[3 * (size_t)i + d] = f[d];
    }
    if (energy) { energy[0] = elj; energy[1] = ecoul; }
    if (virial) *virial = vir;
}

void orc_bond_forces(int n, const float* pos, const float* box3, const orc_ff* ff,
                     int nb, const int* bonds, double* F, double* energy, double* virial) {
    (void)n;
    double e = 0.0, w = 0.0;
    for (int b = 0; b < nb; b++) {
        int i = bonds[3 * b], j = bonds[3 * b + 1], t = bonds[3 * b + 2];
        double d[3];
        for (int k = 0; k < 3; k++) {
            double dd = (double)pos[3 * (size_t)i + k] - (double)pos[3 * (size_t)j + k];
            d[k] = dd - box3[k] * std::rint(dd / box3[k]);
        }
        double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = std::sqrt(rsq);
        double fa = ff->bk2[t] * (ff->br0[t] / r - 1.0);   // src/bonded_forces.cu:75
        for (int k = 0; k < 3; k++) {
            F[3 * (size_t)i + k] += fa * d[k];
            F[3 * (size_t)j + k] -= fa * d[k];
        }
        e += 0.5 * ff->bk2[t] * (r - ff->br0[t]) * (r - ff->br0[t]);
        w += fa * rsq;
    }
    if (energy) *energy = e;
    if (virial) *virial = w;
}

void orc_angle_forces(int n, const float* pos, const float* box3, const orc_ff* ff,
                      int na, const int* angles, double* F, double* energy) {
    (void)n;
    double e = 0.0;
    for (int a = 0; a < na; a++) {
        int i = angles[4 * a], j = angles[4 * a + 1], k = angles[4 * a + 2], t = angles[4 * a + 3];
        double va[3], vb[3];
        for (int d = 0; d < 3; d++) {
            double x = (double)pos[3 * (size_t)i + d] - (double)pos[3 * (size_t)j + d];
            va[d] = x - box3[d] * std::rint(x / box3[d]);
            double y = (double)pos[3 * (size_t)k + d] - (double)pos[3 * (size_t)j + d];
            vb[d] = y - box3[d] * std::rint(y / box3[d]);
        }
        double ra = std::sqrt(va[0] * va[0] + va[1] * va[1] + va[2] * va[2]);
        double rb = std::sqrt(vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]);
        double ab = va[0] * vb[0] + va[1] * vb[1] + va[2] * vb[2];
        double c = ab / (ra * rb);
        c = std::min(std::max(c, -0.9999999), 0.9999999);   // src/bonded_forces.cu:154-155
        double th = std::acos(c);
        double fk = ff->ak2[t] * (th - ff->at0[t]) / std::sqrt(1.0 - c * c) / (ra * rb);  // :165
        // end atom i: F = fk (b - (a.b/|a|^2) a); end k: a<->b; center: minus both (:168-178)
        double Fi[3], Fk[3];
        for (int d = 0; d < 3; d++) {
            Fi[d] = fk * (vb[d] - ab / (ra * ra) * va[d]);
            Fk[d] = fk * (va[d] - ab / (rb * rb) * vb[d]);
            F[3 * (size_t)i + d] += Fi[d];
            F[3 * (size_t)k + d] += Fk[d];
            F[3 * (size_t)j + d] -= Fi[d] + Fk[d];
        }
        e += 0.5 * ff->ak2[t] * (th - ff->at0[t]) * (th - ff->at0[t]);
    }
    if (energy) *energy = e;
}

}  // extern "C"

// ---------------- rigid 3-site water: SETTLE and SHAKE ----------------
namespace {

// Canonical triangle about the centre of mass: O at (0, ra, 0), H at (-+rc, -rb, 0)
// with ra + rb the height of the triangle and mA ra = 2 mB rb.
struct SettleGeom { double mA, mB, M, ra, rb, rc; };

SettleGeom settle_geom(double mA, double mB, double d_oh, double d_hh) {
    SettleGeom g;
    g.mA = mA;
    g.mB = mB;
    g.M = mA + 2.0 * mB;
    g.rc = 0.5 * d_hh;
    const double h = std::sqrt(d_oh * d_oh - g.rc * g.rc);
    g.ra = 2.0 * mB * h / g.M;
    g.rb = h - g.ra;
    return g;
}

inline double mimg(double d, double L) { return d - L * std::rint(d / L); }
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
inline void unit3(double* a) {
    const double r = std::sqrt(dot3(a, a));
    for (int d = 0; d < 3; d++) a[d] /= r;
}

// One molecule.  B0, C0: old H - O vectors; B1, C1: new H - O vectors (the new O
// at the origin).  Out: the corrections of O, H1, H2.
//  1. centre of mass of the new positions; the frame ez = unit(B0 x C0) (normal to
//     the old plane), ex = unit(a1 x ez), ey = ez x ex;
//  2. the corrections lie in the old plane, so each atom keeps its ez coordinate:
//     the canonical triangle is tilted by phi (about ex) and psi (about ey) to match
//     them: sin phi = z_a1 / ra, sin psi = (z_b1 - z_c1) / (2 rc cos phi);
//  3. the turn theta about ez makes the in-plane corrections combinations of the
//     old bond vectors: alpha sin(theta) ... the paper's eqs. (A17)-(A20):
//     sin theta = (alpha gamma - beta sqrt(alpha^2 + beta^2 - gamma^2)) / (alpha^2 + beta^2).
void settle1(const SettleGeom& g, const double* B0, const double* C0, const double* B1, const double* C1,
             double* dA, double* dB, double* dC) {
    double a1[3], b1[3], c1[3];
    for (int d = 0; d < 3; d++) {
        const double com = g.mB * (B1[d] + C1[d]) / g.M;
        a1[d] = -com;
        b1[d] = B1[d] - com;
        c1[d] = C1[d] - com;
    }
    double ez[3], ex[3], ey[3];
    cross3(B0, C0, ez);
    unit3(ez);
    cross3(a1, ez, ex);
    unit3(ex);
    cross3(ez, ex, ey);
    const double xb0 = dot3(B0, ex), yb0 = dot3(B0, ey), xc0 = dot3(C0, ex), yc0 = dot3(C0, ey);
    const double za1 = dot3(a1, ez);
    const double xb1 = dot3(b1, ex), yb1 = dot3(b1, ey), zb1 = dot3(b1, ez);
    const double xc1 = dot3(c1, ex), yc1 = dot3(c1, ey), zc1 = dot3(c1, ez);
    const double sinphi = za1 / g.ra;
    const double cosphi = std::sqrt(std::max(0.0, 1.0 - sinphi * sinphi));
    const double sinpsi = (zb1 - zc1) / (2.0 * g.rc * cosphi);
    const double cospsi = std::sqrt(std::max(0.0, 1.0 - sinpsi * sinpsi));
    const double ya2 = g.ra * cosphi;
    const double xb2 = -g.rc * cospsi;
    const double t1 = -g.rb * cosphi, t2 = g.rc * sinpsi * sinphi;
    const double yb2 = t1 - t2, yc2 = t1 + t2;
    const double alpha = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
    const double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
    const double gamma = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
    const double ab2 = alpha * alpha + beta * beta;
    const double sinth = (alpha * gamma - beta * std::sqrt(std::max(0.0, ab2 - gamma * gamma))) / ab2;
    const double costh = std::sqrt(std::max(0.0, 1.0 - sinth * sinth));
    const double a3[3] = {-ya2 * sinth, ya2 * costh, za1};
    const double b3[3] = {xb2 * costh - yb2 * sinth, xb2 * sinth + yb2 * costh, zb1};
    const double c3[3] = {-xb2 * costh - yc2 * sinth, -xb2 * sinth + yc2 * costh, zc1};
    for (int d = 0; d < 3; d++) {
        dA[d] = a3[0] * ex[d] + a3[1] * ey[d] + a3[2] * ez[d] - a1[d];
        dB[d] = b3[0] * ex[d] + b3[1] * ey[d] + b3[2] * ez[d] - b1[d];
        dC[d] = c3[0] * ex[d] + c3[1] * ey[d] + c3[2] * ez[d] - c1[d];
    }
}

// the O -> H vectors of molecule m at positions x (minimum image)
void oh_vectors(const int* mol, int m, const double* x, const double* L, double* B, double* C) {
    const int a = mol[3 * m], b = mol[3 * m + 1], c = mol[3 * m + 2];
    for (int d = 0; d < 3; d++) {
        B[d] = mimg(x[3 * (size_t)b + d] - x[3 * (size_t)a + d], L[d]);
        C[d] = mimg(x[3 * (size_t)c + d] - x[3 * (size_t)a + d], L[d]);
    }
}

}  // namespace

extern "C" void orc_settle(int nmol, const int* mol, const double* mass, const float* box3, double d_oh,
                           double d_hh, const double* x0, double* x1) {
    const double L[3] = {box3[0], box3[1], box3[2]};
#pragma omp parallel for schedule(static)
    for (int m = 0; m < nmol; m++) {
        const int a = mol[3 * m], b = mol[3 * m + 1], c = mol[3 * m + 2];
        const SettleGeom g = settle_geom(mass[a], mass[b], d_oh, d_hh);
        double B0[3], C0[3], B1[3], C1[3], dA[3], dB[3], dC[3];
        oh_vectors(mol, m, x0, L, B0, C0);
        oh_vectors(mol, m, x1, L, B1, C1);
        settle1(g, B0, C0, B1, C1, dA, dB, dC);
        for (int d = 0; d < 3; d++) {
            x1[3 * (size_t)a + d] += dA[d];
            x1[3 * (size_t)b + d] += dB[d];
            x1[3 * (size_t)c + d] += dC[d];
        }
    }
}

extern "C" int orc_shake(int nmol, const int* mol, const double* mass, const float* box3, double d_oh,
                         double d_hh, const double* x0, double* x1, double tol, int maxit) {
    const double L[3] = {box3[0], box3[1], box3[2]};
    int worst = 0;
#pragma omp parallel for schedule(static) reduction(max : worst)
    for (int m = 0; m < nmol; m++) {
        const int id[3] = {mol[3 * m], mol[3 * m + 1], mol[3 * m + 2]};
        double X0[3][3] = {{0, 0, 0}}, X[3][3] = {{0, 0, 0}}, Xs[3][3];
        oh_vectors(mol, m, x0, L, X0[1], X0[2]);
        oh_vectors(mol, m, x1, L, X[1], X[2]);
        std::memcpy(Xs, X, sizeof X);
        const double im[3] = {1.0 / mass[id[0]], 1.0 / mass[id[1]], 1.0 / mass[id[2]]};
        const int pi[3] = {0, 0, 1}, pj[3] = {1, 2, 2};
        const double d2[3] = {d_oh * d_oh, d_oh * d_oh, d_hh * d_hh};
        int used = -1;
        for (int it = 1; it <= maxit && used < 0; it++) {
            bool done = true;
            for (int k = 0; k < 3; k++) {
                const int i = pi[k], j = pj[k];
                double s[3], r0[3];
                for (int d = 0; d < 3; d++) {
                    s[d] = X[i][d] - X[j][d];
                    r0[d] = X0[i][d] - X0[j][d];
                }
                const double diff = d2[k] - dot3(s, s);
                if (std::fabs(diff) > tol * d2[k]) {
                    done = false;
                    const double gk = diff / (2.0 * dot3(s, r0) * (im[i] + im[j]));
                    for (int d = 0; d < 3; d++) {
                        X[i][d] += gk * r0[d] * im[i];
                        X[j][d] -= gk * r0[d] * im[j];
                    }
                }
            }
            if (done) used = it;
        }
        for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) x1[3 * (size_t)id[a] + d] += X[a][d] - Xs[a][d];
        worst = std::max(worst, used < 0 ? (1 << 30) : used);
    }
    return worst >= (1 << 30) ? -1 : worst;
}

// ---------------- FP64 MD driver ----------------
namespace {

struct MDSys {
    int n;
    const int* type;
    const double* charge;
    const double* mass;
    double L[3];
    const orc_ff* ff;
    int nb, na;
    const int* bonds;
    const int* angles;
    const int* excl;
    double rl;
    orc_pme pme{};   // order 0: no reciprocal part
    std::vector<int> counts, nbrs;
    int cap = 0;
    std::vector<double> pos_build;
};

// FP64 list with FP64 predicate r^2 < rl^2 (the oracle trajectory is pure FP64).
void md_build(MDSys& s, const double* pos) {
    int n = s.n;
    std::vector<std::vector<int>> L(n);
    Cells c;
    double rl2 = s.rl * s.rl;
    auto test = [&](int i, int j) {
        if (j == i || excluded(s.excl, i, j)) return;
        double r2 = 0;
        for (int d = 0; d < 3; d++) {
            double dd = pos[3 * (size_t)i + d] - pos[3 * (size_t)j + d];
            dd -= s.L[d] * std::rint(dd / s.L[d]);
            r2 += dd * dd;
        }
        if (r2 < rl2) L[i].push_back(j);
    };
    if (build_cells(n, pos, s.L, s.rl, c)) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; i++) for_cell_neighbors(c, pos + 3 * (size_t)i, s.L, [&](int j) { test(i, j); });
    } else {
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) test(i, j);
    }
    int m = 0;
    for (auto& v : L) m = std::max(m, (int)v.size());
    s.cap = m;
    s.counts.assign(n, 0);
    s.nbrs.assign((size_t)n * std::max(m, 1), 0);
    finish_lists(n, std::max(m, 1), L, s.counts.data(), s.nbrs.data(), &m);
    s.cap = std::max(s.cap, 1);
    s.pos_build.assign(pos, pos + 3 * (size_t)n);
}

bool md_need_rebuild(const MDSys& s, const double* pos, double skin) {
    double lim = skin * skin / 4.0;  // src/neighborlist.cu:36,174
    for (int i = 0; i < s.n; i++) {
        double r2 = 0;
        for (int d = 0; d < 3; d++) {
            double dd = pos[3 * (size_t)i + d] - s.pos_build[3 * (size_t)i + d];
            dd -= s.L[d] * std::rint(dd / s.L[d]);
            r2 += dd * dd;
        }
        if (r2 >= lim) return true;
    }
    return false;
}

// Forces (FP64 positions), returns potential energy.
// Forces and potential energy; *vir (if given) = W, the sum over pairs and
// bonds of F.r (each pair once).
double md_forces(MDSys& s, const double* pos, double* F, double* vir = nullptr) {
    const orc_ff* ff = s.ff;
    int n = s.n, nt = ff->ntype;
    CoulConst cc = coul_const(ff);
    double rc2 = ff->rc * ff->rc;
    double epot = 0.0, w = 0.0;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : epot, w)
    for (int i = 0; i < n; i++) {
        double f[3] = {0, 0, 0};
        for (int k = 0; k < s.counts[i]; k++) {
            int j = s.nbrs[(size_t)i * s.cap + k];
            double d[3];
            for (int q = 0; q < 3; q++) {
                double dd = pos[3 * (size_t)i + q] - pos[3 * (size_t)j + q];
                d[q] = dd - s.L[q] * std::rint(dd / s.L[q]);
            }
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (rsq > rc2) continue;
            double r = std::sqrt(rsq);
            int tp = s.type[i] * nt + s.type[j];
            double qq = ff->qfac * s.charge[i] * s.charge[j];
            double fr, ur;
            coul_terms(ff, cc, r, &fr, &ur);
            double ffv = qq * fr + lj_force(ff->lj1[tp], ff->lj2[tp], rsq);
            for (int q = 0; q < 3; q++) f[q] += ffv * d[q];
            epot += 0.5 * (qq * ur + lj_energy(ff->lj1[tp], ff->lj2[tp], rsq) - lj_energy(ff->lj1[tp], ff->lj2[tp], rc2));
            w += 0.5 * ffv * rsq;
        }
        for (int q = 0; q < 3; q++) F[3 * (size_t)i + q] = f[q];
    }
    // bonded terms in FP64 positions
    for (int b = 0; b < s.nb; b++) {
        int i = s.bonds[3 * b], j = s.bonds[3 * b + 1], t = s.bonds[3 * b + 2];
        double d[3];
        for (int q = 0; q < 3; q++) {
            double dd = pos[3 * (size_t)i + q] - pos[3 * (size_t)j + q];
            d[q] = dd - s.L[q] * std::rint(dd / s.L[q]);
        }
        double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double fa = ff->bk2[t] * (ff->br0[t] / r - 1.0);
        for (int q = 0; q < 3; q++) { F[3 * (size_t)i + q] += fa * d[q]; F[3 * (size_t)j + q] -= fa * d[q]; }
        epot += 0.5 * ff->bk2[t] * (r - ff->br0[t]) * (r - ff->br0[t]);
        w += fa * r * r;
    }
    for (int a = 0; a < s.na; a++) {
        int i = s.angles[4 * a], j = s.angles[4 * a + 1], k = s.angles[4 * a + 2], t = s.angles[4 * a + 3];
        double va[3], vb[3];
        for (int q = 0; q < 3; q++) {
            double x = pos[3 * (size_t)i + q] - pos[3 * (size_t)j + q];
            va[q] = x - s.L[q] * std::rint(x / s.L[q]);
            double y = pos[3 * (size_t)k + q] - pos[3 * (size_t)j + q];
            vb[q] = y - s.L[q] * std::rint(y / s.L[q]);
        }
        double ra = std::sqrt(va[0] * va[0] + va[1] * va[1] + va[2] * va[2]);
        double rb = std::sqrt(vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]);
        double ab = va[0] * vb[0] + va[1] * vb[1] + va[2] * vb[2];
        double c = std::min(std::max(ab / (ra * rb), -0.9999999), 0.9999999);
        double th = std::acos(c);
        double fk = ff->ak2[t] * (th - ff->at0[t]) / std::sqrt(1.0 - c * c) / (ra * rb);
        for (int q = 0; q < 3; q++) {
            double Fi = fk * (vb[q] - ab / (ra * ra) * va[q]);
            double Fk = fk * (va[q] - ab / (rb * rb) * vb[q]);
            F[3 * (size_t)i + q] += Fi;
            F[3 * (size_t)k + q] += Fk;
            F[3 * (size_t)j + q] -= Fi + Fk;
        }
        epot += 0.5 * ff->ak2[t] * (th - ff->at0[t]) * (th - ff->at0[t]);
    }
    // Ewald reciprocal part, excluded-pair correction and self term (md_pme.cpp)
    if (s.pme.order > 0) {
        const float b3[3] = {(float)s.L[0], (float)s.L[1], (float)s.L[2]};
        double w1 = 0.0, w2 = 0.0, w3 = 0.0;
        epot += orc_pme_recip(&s.pme, n, pos, s.charge, b3, ff->qfac, F, &w1);
        if (s.excl) epot += orc_pme_excl(n, pos, s.charge, b3, s.excl, s.pme.beta, ff->qfac, F, &w2);
        epot += orc_ewald_self(n, s.charge, b3, s.pme.beta, ff->qfac, &w3);
        w += w1 + w2 + w3;
    }
    if (vir) *vir = w;
    return epot;
}

double kinetic2(int n, const double* vel, const double* mass) {   // sum m_int v^2 (kJ/mol)
    double s = 0.0;
    for (int i = 0; i < n; i++) {
        double m = 1e4 * mass[i];   // src/simulation.cu:177
        s += m * (vel[3 * (size_t)i] * vel[3 * (size_t)i] + vel[3 * (size_t)i + 1] * vel[3 * (size_t)i + 1] +
                  vel[3 * (size_t)i + 2] * vel[3 * (size_t)i + 2]);
    }
    return s;
}

inline void wrap(double* x, const double* L) {   // src/nve_leapfrog.cu:50-55 (box centred at 0)
    for (int d = 0; d < 3; d++) x[d] -= L[d] * std::trunc(2.0 * x[d] / L[d]);
}

}  // namespace

extern "C" int orc_md_run(int n, double* pos, double* vel, const int* type, const double* charge,
                          const double* mass, const float* box3, const orc_ff* ff, const orc_md* md,
                          int nb, const int* bonds, int na, const int* angles, const int* excl,
                          int nsteps, double* ke, double* pe, double* temp, double* conserved,
                          double* thermo_state, double* volume, double* box_final) {
    MDSys s;
    s.n = n; s.type = type; s.charge = charge; s.mass = mass;
    for (int d = 0; d < 3; d++) s.L[d] = box3[d];
    s.ff = ff; s.nb = nb; s.na = na; s.bonds = bonds; s.angles = angles; s.excl = excl;
    s.rl = ff->rc + md->skin;
    const double dt = md->dt;
    // src/integrator.cu:155, less the three constraints of each rigid molecule
    const double Nf = 3.0 * n - 3.0 - 3.0 * (md->nrigid > 0 ? md->nrigid : 0);
    if (md->nrigid > 0 && md->integrator != ORC_INT_LEAPFROG) return -1;   // SETTLE with leapfrog only
    if (md->pme_order > 0) {   // damped Coulomb (the reciprocal part uses the box of each step)
        if (ff->coulomb != ORC_COUL_DAMPED) return -2;
        s.pme.order = md->pme_order;
        for (int d = 0; d < 3; d++) s.pme.K[d] = md->pme_K[d];
        s.pme.beta = ff->alpha;
    }
    std::vector<double> x_old(md->nrigid > 0 ? 3 * (size_t)n : 0);
    std::vector<double> F(3 * (size_t)n);
    std::vector<double> im(n);
    for (int i = 0; i < n; i++) im[i] = 1.0 / (1e4 * mass[i]);
    md_build(s, pos);
    double W = 0.0;
    double U = md_forces(s, pos, F.data(), &W);
    int rebuilds = 1;
    const bool baro = md->barostat == ORC_BARO_BERENDSEN && md->integrator == ORC_INT_LEAPFROG;
    // thermostat state
    double xi = thermo_state ? thermo_state[0] : 0.0, vxi = thermo_state ? thermo_state[1] : 0.0;
    const double kT0 = KB_REF * md->T0;
    const double Q = Nf * kT0 * md->tau * md->tau;    // NH mass, Q = Nf kB T0 tau^2
    // Berendsen leapfrog keeps the previous step's temperature (src/integrator.cu:77,146-155)
    double T_prev = md->T0;

    for (int step = 0; step < nsteps; step++) {
        if (md->integrator == ORC_INT_LEAPFROG) {
            double lam = 1.0;
            if (md->thermostat == ORC_TH_BERENDSEN)   // src/integrator.cu:146-147
                lam = std::sqrt(1.0 + dt / md->tau * (md->T0 / T_prev - 1.0));
            if (md->nrigid > 0) std::memcpy(x_old.data(), pos, 3 * (size_t)n * sizeof(double));
            for (int i = 0; i < n; i++)
                for (int d = 0; d < 3; d++) {
                    size_t k = 3 * (size_t)i + d;
                    vel[k] = (vel[k] + F[k] * dt * im[i]) * lam;   // src/nvt_leapfrog_ber.cu:39-41
                    pos[k] += vel[k] * dt;
                }
            if (md->nrigid > 0) {   // constrained positions; the velocities they imply
                orc_settle(md->nrigid, md->rigid, mass, box3, md->d_oh, md->d_hh, x_old.data(), pos);
                for (int m = 0; m < md->nrigid; m++)
                    for (int a = 0; a < 3; a++) {
                        const size_t k = 3 * (size_t)md->rigid[3 * m + a];
                        for (int d = 0; d < 3; d++) vel[k + d] = (pos[k + d] - x_old[k + d]) / dt;
                    }
            }
            for (int i = 0; i < n; i++) wrap(pos + 3 * (size_t)i, s.L);
            double k2 = kinetic2(n, vel, mass);
            T_prev = k2 / Nf / KB_REF;                    // src/integrator.cu:153-155
            // step statistics: KE of v(t+dt/2) and the potential energy at x(t),
            // i.e. of the forces this step used
            ke[step] = 0.5 * k2; pe[step] = U; temp[step] = T_prev;
            conserved[step] = 0.5 * k2 + U;
            if (volume) volume[step] = s.L[0] * s.L[1] * s.L[2];
            if (baro) {   // Berendsen barostat: P of this step, then positions and box scaled
                const double V = s.L[0] * s.L[1] * s.L[2];
                const double P = (k2 + W) / (3.0 * V) * ORC_BAR_PER_KJMOL_A3;
                const double mu = std::cbrt(1.0 - md->compress * dt / md->tau_p * (md->P0 - P));
                for (size_t k = 0; k < 3 * (size_t)n; k++) pos[k] *= mu;
                for (int d = 0; d < 3; d++) s.L[d] *= mu;
                md_build(s, pos);   // the scaled pair distances: a list per step
                rebuilds++;
            } else if (md_need_rebuild(s, pos, md->skin)) {
                md_build(s, pos);
                rebuilds++;
            }
            U = md_forces(s, pos, F.data(), &W);
        } else {
            // velocity Verlet (textbook; the reference's vverlet_nve_pre double-kicks, src/nve_vverlet.cu:49-54)
            auto nh_half = [&](double& K) {   // MTK chain-of-one half step (Martyna, Tuckerman, Klein 1996)
                double G = (2.0 * K - Nf * kT0) / Q;
                vxi += 0.25 * dt * G;
                double sc = std::exp(-0.5 * dt * vxi);
                for (size_t k = 0; k < 3 * (size_t)n; k++) vel[k] *= sc;
                K *= sc * sc;
                xi += 0.5 * dt * vxi;
                G = (2.0 * K - Nf * kT0) / Q;
                vxi += 0.25 * dt * G;
            };
            double K = 0.5 * kinetic2(n, vel, mass);
            if (md->thermostat == ORC_TH_NOSEHOOVER) nh_half(K);
            for (int i = 0; i < n; i++)
                for (int d = 0; d < 3; d++) {
                    size_t k = 3 * (size_t)i + d;
                    vel[k] += 0.5 * dt * F[k] * im[i];
                    pos[k] += vel[k] * dt;
                }
            for (int i = 0; i < n; i++) wrap(pos + 3 * (size_t)i, s.L);
            if (md_need_rebuild(s, pos, md->skin)) { md_build(s, pos); rebuilds++; }
            U = md_forces(s, pos, F.data());
            for (int i = 0; i < n; i++)
                for (int d = 0; d < 3; d++) vel[3 * (size_t)i + d] += 0.5 * dt * F[3 * (size_t)i + d] * im[i];
            K = 0.5 * kinetic2(n, vel, mass);
            if (md->thermostat == ORC_TH_NOSEHOOVER) nh_half(K);
            if (md->thermostat == ORC_TH_BERENDSEN) {   // current-temperature Berendsen scaling
                double T = 2.0 * K / Nf / KB_REF;
                double lam = std::sqrt(1.0 + dt / md->tau * (md->T0 / T - 1.0));
                for (size_t k = 0; k < 3 * (size_t)n; k++) vel[k] *= lam;
                K *= lam * lam;
            }
            ke[step] = K; pe[step] = U; temp[step] = 2.0 * K / Nf / KB_REF;
            conserved[step] = K + U;
            if (md->thermostat == ORC_TH_NOSEHOOVER) conserved[step] += 0.5 * Q * vxi * vxi + Nf * kT0 * xi;
        }
    }
    if (thermo_state) { thermo_state[0] = xi; thermo_state[1] = vxi; }
    if (box_final)
        for (int d = 0; d < 3; d++) box_final[d] = s.L[d];
    return rebuilds;
}

// ---------------- force matching ----------------
extern "C" double orc_fm_loss_grad(int nframes, int n, const float* pos, const int* type,
                                   const double* charge, const float* box3, const orc_ff* ff,
                                   double qscale, const double* target,
                                   double* g_lj1, double* g_lj2, double* g_qscale) {
    const int nt = ff->ntype;
    CoulConst cc = coul_const(ff);
    float rc2, rl2, sk2, linv[3];
    orc_canonical_thresholds((float)ff->rc, 0.5f, box3, &rc2, &rl2, &sk2, linv);
    const double M = 3.0 * n * nframes;
    double loss = 0.0, gq = 0.0;
    std::vector<double> g1(nt * nt, 0.0), g2(nt * nt, 0.0);
    std::vector<int> counts(n);
    for (int f = 0; f < nframes; f++) {
        const float* P = pos + (size_t)f * n * 3;
        const double* T = target + (size_t)f * n * 3;
        int cap = 1;
        {   // size the list
            std::vector<int> tmp(n);
            std::vector<int> dummy((size_t)n);
            cap = std::max(1, orc_nlist_cells(n, P, box3, rl2, nullptr, 1, counts.data(), dummy.data()));
        }
        std::vector<int> nb((size_t)n * cap);
        orc_nlist_cells(n, P, box3, rl2, nullptr, cap, counts.data(), nb.data());
        std::vector<double> F(3 * (size_t)n);
        // pass 1: forces with scaled charges
        std::vector<double> q(n);
        for (int i = 0; i < n; i++) q[i] = qscale * charge[i];
        orc_pair_forces(n, P, type, q.data(), box3, ff, rc2, cap, counts.data(), nb.data(), F.data(), nullptr, nullptr);
        std::vector<double> R(3 * (size_t)n);
        for (size_t k = 0; k < 3 * (size_t)n; k++) {
            R[k] = F[k] - T[k];
            loss += R[k] * R[k];
        }
        // pass 2: adjoint over ordered pairs: dL/dtheta = 2/M sum_i sum_j (R_i.d_ij) dff_ij/dtheta
#pragma omp parallel
        {
            std::vector<double> l1(nt * nt, 0.0), l2(nt * nt, 0.0);
            double lq = 0.0;
#pragma omp for schedule(dynamic, 32)
            for (int i = 0; i < n; i++) {
                const float* xi = P + 3 * (size_t)i;
                for (int k = 0; k < counts[i]; k++) {
                    int j = nb[(size_t)i * cap + k];
                    const float* xj = P + 3 * (size_t)j;
                    float d32[3], s[3];
                    for (int d = 0; d < 3; d++) {
                        float dd = xi[d] - xj[d];
                        s[d] = shift32(dd, linv[d]);
                        d32[d] = fmaf(-box3[d], s[d], dd);
                    }
                    if (rsq32(d32[0], d32[1], d32[2]) > rc2) continue;
                    double d64[3];
                    for (int d = 0; d < 3; d++) d64[d] = ((double)xi[d] - (double)xj[d]) - (double)box3[d] * (double)s[d];
                    double rsq = d64[0] * d64[0] + d64[1] * d64[1] + d64[2] * d64[2];
                    double r = std::sqrt(rsq);
                    double g = R[3 * (size_t)i] * d64[0] + R[3 * (size_t)i + 1] * d64[1] + R[3 * (size_t)i + 2] * d64[2];
                    double ri2 = 1.0 / rsq, ri8 = ri2 * ri2 * ri2 * ri2;
                    int tp = type[i] * nt + type[j];
                    l1[tp] += g * ri8 * ri2 * ri2 * ri2;   // d ff / d lj1 = r^-14
                    l2[tp] -= g * ri8;                     // d ff / d lj2 = -r^-8
                    double fr;
                    coul_terms(ff, cc, r, &fr, nullptr);
                    lq += g * 2.0 * qscale * ff->qfac * charge[i] * charge[j] * fr;   // d ff / d qscale
                }
            }
#pragma omp critical
            {
                for (int t = 0; t < nt * nt; t++) { g1[t] += l1[t]; g2[t] += l2[t]; }
                gq += lq;
            }
        }
    }
    for (int t = 0; t < nt * nt; t++) {
        if (g_lj1) g_lj1[t] = 2.0 / M * g1[t];
        if (g_lj2) g_lj2[t] = 2.0 / M * g2[t];
    }
    if (g_qscale) *g_qscale = 2.0 / M * gq;
    return loss / M;
}
