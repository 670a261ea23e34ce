/*
 * md_oracle.h -- FP64 CPU restatement of the md_gpu reference MD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product
 * (md_gpu_amd/, libmdgpu.so) never links, loads or calls anything here.
 *
 * What it restates (reference files, read as text):
 *   pair forces      src/pair_forces.cu:34-204   (RF + LJ, Wolf + LJ)
 *   force constants  src/forces.cu:47-111, src/simulation.cu:184-205, :86, :94
 *   neighbor list    src/neighborlist.cu:35-36, :64-99, :157-176, :231-297
 *   bonds / angles   src/bonded_forces.cu:34-192, src/simulation.cu:225-244
 *   integrators      src/nve_leapfrog.cu:31-59, src/nvt_leapfrog_ber.cu:29-80,
 *                    src/integrator.cu:131-156, src/nve_vverlet.cu:31-78
 *   wrap / units     src/nve_leapfrog.cu:50-55, src/simulation.cu:177
 * Parts with no reference arithmetic (energies, Nose-Hoover, force-matching
 * gradient) follow the textbook forms cited at each function.
 *
 * Pinning status of every function is listed in the header comment of
 * md_oracle.cpp; tests/test_oracle_*.py hold the checks.
 */
#ifndef MD_ORACLE_H
#define MD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_TYPES 8

/* Coulomb kinds, matching the reference's lkspace switch (src/simulation.cu:82-95)
 * plus an explicit "none" used for uncharged systems. */
enum { ORC_COUL_RF = 0, ORC_COUL_WOLF = 1, ORC_COUL_DAMPED = 2, ORC_COUL_NONE = 3 };

typedef struct {
    int    ntype;
    double lj1[ORC_MAX_TYPES * ORC_MAX_TYPES];   /* 48 eps sig^12 (or 12 A), src/simulation.cu:191,200 */
    double lj2[ORC_MAX_TYPES * ORC_MAX_TYPES];   /* 24 eps sig^6  (or 6 B),  src/simulation.cu:194,202 */
    int    coulomb;                               /* ORC_COUL_* */
    double rc;                                    /* cutoff (A) */
    double alpha;                                 /* Wolf / damped alpha (1/A) */
    double eps_rf;                                /* reaction-field dielectric */
    double qfac;                                  /* 1389.35427611722, src/forces.cu:47 */
    int    erfc_exact;                            /* 0: A&S 7.1.26 polynomial (reference); 1: libm erfc */
    int    nbtype;
    double bk2[8], br0[8];                        /* bnd1 = 2k, bnd2 = r0 (src/simulation.cu:226-227) */
    int    natype;
    double ak2[8], at0[8];                        /* ang1 = 2k, ang2 = theta0 rad (src/simulation.cu:241-243) */
} orc_ff;

/* Canonical FP32 thresholds (shared definition with the HIP engine):
 *   rc2   = (float)rc * (float)rc                               src/simulation.cu:96
 *   rl2   = (float)((skin + sqrt((double)rc2))^2)               src/neighborlist.cu:35
 *   skin2q= skin*skin/4 in float                                src/neighborlist.cu:36
 *   Linv  = (float)(1.0/(double)L)                              src/simulation.cu:290-292 */
void orc_canonical_thresholds(float rc, float skin, const float* box3,
                              float* rc2, float* rl2, float* skin2q, float* linv3);

/* Canonical FP32 squared minimum-image distance between two float positions
 * (the predicate the neighbor list and the cutoff test use on both sides). */
float orc_rsq32(const float* xi, const float* xj, const float* box3, const float* linv3);

/* Full neighbor lists, src/neighborlist.cu:231-297 semantics:
 *   j in N(i)  <=>  j != i, j not in excl(i), rsq32(i,j) < rl2.
 * excl: n*4 ints, -1 = empty slot (1-2 bonded partners, src/neighborlist.cu:64-99).
 * Lists are written sorted ascending into nbrs[i*cap + k]; counts[i] is the
 * full count even if it exceeds cap.  Returns max count. */
int orc_nlist_brute(int n, const float* pos, const float* box3, float rl2,
                    const int* excl, int cap, int* counts, int* nbrs);
/* Same sets via a cell list (O(N)); must equal orc_nlist_brute. */
int orc_nlist_cells(int n, const float* pos, const float* box3, float rl2,
                    const int* excl, int cap, int* counts, int* nbrs);

/* Pair forces over full lists (src/pair_forces.cu:133-204 / :34-131), FP64.
 * Pair included iff rsq32 <= rc2 (src/pair_forces.cu:95,195; the reference
 * zeroes ff when rsq > rcutsq).  F[i*3..] = sum_j ff*d_ij.
 * energy[0] = LJ (shifted to 0 at rc), energy[1] = Coulomb, virial = sum ff r^2
 * over unordered pairs (both halves counted once).  Pass energy/virial NULL to skip. */
void orc_pair_forces(int n, const float* pos, const int* type, const double* charge,
                     const float* box3, const orc_ff* ff, float rc2,
                     int cap, const int* counts, const int* nbrs,
                     double* F, double* energy, double* virial);

/* Harmonic bonds (src/bonded_forces.cu:34-95), E = k (r-r0)^2 with bnd1 = 2k.
 * bonds: nb*3 ints (i, j, type), 0-based.  Adds into F.  energy/virial may be NULL. */
void orc_bond_forces(int n, const float* pos, const float* box3, const orc_ff* ff,
                     int nb, const int* bonds, double* F, double* energy, double* virial);

/* Harmonic angles (src/bonded_forces.cu:98-192), E = k (theta-theta0)^2 with ang1 = 2k.
 * angles: na*4 ints (i, center j, k, type), 0-based.  Adds into F. */
void orc_angle_forces(int n, const float* pos, const float* box3, const orc_ff* ff,
                      int na, const int* angles, double* F, double* energy);

/* ---- MD trajectories in FP64 (positions and velocities double) ---- */
enum { ORC_INT_LEAPFROG = 0, ORC_INT_VVERLET = 1 };
enum { ORC_TH_NONE = 0, ORC_TH_BERENDSEN = 1, ORC_TH_NOSEHOOVER = 2 };
/* Berendsen barostat (Berendsen et al., J. Chem. Phys. 81, 3684 (1984)), isotropic:
 * after each leapfrog step P = (sum m v^2 + W) / (3 V) with W = sum over pairs and
 * bonds of F.r at the forces of the step (angles carry no virial, as in the
 * reference's bonded kernels), mu = cbrt(1 - compress dt / tau_p (P0 - P)),
 * positions and box scaled by mu.  The reference declares press_0 and pcons
 * (src/simulation.h:113-116) and an NPT kernel without box scaling
 * (src/npt_leapfrog_ber.cu:28-123, its call commented out at
 * src/integrator.cu:157-164). */
enum { ORC_BARO_NONE = 0, ORC_BARO_BERENDSEN = 1 };
#define ORC_BAR_PER_KJMOL_A3 16605.390671738467   /* 1 kJ/mol/A^3 in bar */

typedef struct {
    int    integrator;      /* ORC_INT_* */
    int    thermostat;      /* ORC_TH_* */
    double dt;              /* fs */
    double T0;              /* K */
    double tau;             /* fs (Berendsen tau = 1/tco, src/integrator.cu:146; NH period) */
    double skin;            /* A */
    int    barostat;        /* ORC_BARO_*: leapfrog only */
    double P0;              /* bar */
    double tau_p;           /* fs */
    double compress;        /* isothermal compressibility, 1/bar */
    int    nrigid;          /* rigid 3-site molecules (leapfrog only): SETTLE after each position update */
    const int* rigid;       /* nrigid x 3 atom indices (O, H1, H2) */
    double d_oh, d_hh;      /* their O-H and H-H distances (A) */
    int    pme_order;       /* > 0 (damped Coulomb only): add the smooth-PME reciprocal part, the */
    int    pme_K[3];        /* excluded-pair correction and the self term at beta = ff->alpha, in
                               the box of each step (pressure coupling rescales it) */
} orc_md;

/* Stats per step (length nsteps): ke, pe (lj+coul+bond+angle), temperature,
 * conserved quantity (H for NVE; H + NH terms for Nose-Hoover).
 * masses: per atom (g/mol).  State arrays are updated in place.
 * Leapfrog: vel holds v(t-dt/2) on entry and exit (reference convention);
 * velocity Verlet: vel holds full-step v(t). */
int orc_md_run(int n, double* pos, double* vel, const int* type, const double* charge,
               const double* mass, const float* box3, const orc_ff* ff, const orc_md* md,
               int nb, const int* bonds, int na, const int* angles, const int* excl,
               int nsteps, double* ke, double* pe, double* temp, double* conserved,
               double* thermo_state /* [xi, v_xi] in/out, may be NULL */,
               double* volume /* per step: the box volume of the step's forces, may be NULL */,
               double* box_final /* [3] the box after the run, may be NULL */);

/* ---- rigid 3-site water ----
 * SETTLE (Miyamoto and Kollman, J. Comput. Chem. 13, 952 (1992)), restated from the
 * paper: the reference's constraint code is uncompiled pseudo-code (src/settle.c,
 * src/settle.cu:1-178, src/xshake.cu:15-82).  mol: nmol x 3 atom indices (O, H1,
 * H2; the two H of equal mass); x0: positions before the step (satisfying the
 * constraints); x1: after the unconstrained update, constrained in place.  The
 * vectors from O to H are taken by minimum image and the corrections added to
 * each atom's own coordinates, so no atom changes its periodic image. */
void orc_settle(int nmol, const int* mol, const double* mass, const float* box3, double d_oh, double d_hh,
                const double* x0, double* x1);
/* SHAKE (Ryckaert, Ciccotti and Berendsen, J. Comput. Phys. 23, 327 (1977)) on the
 * same three distances per molecule, iterated until |d^2 - r^2| < tol d^2 for all:
 * the independent check of orc_settle (both give the positions whose corrections
 * lie along the old bond vectors).  Returns the most iterations a molecule used,
 * -1 if one did not converge within maxit. */
int orc_shake(int nmol, const int* mol, const double* mass, const float* box3, double d_oh, double d_hh,
              const double* x0, double* x1, double tol, int maxit);

/* ---- Ewald reciprocal space (md_pme.cpp) ----
 * Smooth particle-mesh Ewald (Essmann et al., J. Chem. Phys. 103, 8577 (1995)),
 * restated from the paper: the reference vendors GROMACS' PME (src/pme.c) without
 * building it, and its lkspace 2 path has only the damped real-space term
 * (src/forces.cu:97-110).  Positions in the box centred at 0; u = K (x / L + 1/2). */
#define ORC_PME_MAX_ORDER 12
typedef struct {
    int    K[3];            /* grid points per box edge */
    int    order;           /* B-spline order p (2..12) */
    double beta;            /* Ewald splitting parameter (1/A), the damped term's alpha */
} orc_pme;
/* Unnormalized 3D DFT in place of a K0 x K1 x K2 complex array (interleaved re, im;
 * last index fastest); sign -1 forward, +1 backward. */
void orc_fft3(double* data, const int* K, int sign);
/* th[j] = M_p(w + j), dth[j] = M_p'(w + j), j = 0..p-1 (cardinal B-spline of order p). */
void orc_bspline(int order, double w, double* th, double* dth);
/* SPME reciprocal energy (kJ/mol); adds the forces to F (n x 3, may be NULL);
 * *W = trace of the reciprocal virial (may be NULL). */
double orc_pme_recip(const orc_pme* pm, int n, const double* x, const double* q, const float* box3,
                     double qfac, double* F, double* W);
/* The Ewald reciprocal sum over all m = (a/L0, b/L1, c/L2), |a|,|b|,|c| <= kmax, m != 0,
 * by structure factors: the limit SPME converges to.  Same outputs. */
double orc_ewald_recip(int n, const double* x, const double* q, const float* box3, double beta, double qfac,
                       int kmax, double* F, double* W);
/* Excluded pairs (excl: n x 4 partner indices, -1 empty, each pair listed from both
 * atoms): -qfac qi qj erf(beta r)/r per pair at the minimum image, removing what the
 * reciprocal sum adds for pairs the real-space sum leaves out.  Same outputs. */
double orc_pme_excl(int n, const double* x, const double* q, const float* box3, const int* excl, double beta,
                    double qfac, double* F, double* W);
/* Self term -qfac beta/sqrt(pi) sum q^2 plus the neutralizing-background term
 * -qfac pi (sum q)^2 / (2 V beta^2); *W = its virial (3 x the background term). */
double orc_ewald_self(int n, const double* q, const float* box3, double beta, double qfac, double* W);

/* ---- force matching (config: NaCl, 7 parameters) ----
 * loss = mean over frames*atoms*3 of (F(theta) - Ftarget)^2 with charges
 * q_i = qscale * charge[i].  Gradient w.r.t. lj1[a*nt+b], lj2[a*nt+b] (each
 * ordered entry separately) and qscale.  FP64 analytic adjoint. */
double orc_fm_loss_grad(int nframes, int n, const float* pos /* frames*n*3 */,
                        const int* type, const double* charge, const float* box3,
                        const orc_ff* ff, double qscale, const double* target /* frames*n*3 */,
                        double* g_lj1, double* g_lj2, double* g_qscale);

/* OpenMP threads the oracle uses (omp_get_max_threads). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
