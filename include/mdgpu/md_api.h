/*
 * md_api.h -- C ABI of the MI355X MD step engine (libmdgpu.so).
 *
 * Plain C types only: host pointers and sizes, no HIP or torch types.  Every
 * call is synchronous at return unless documented otherwise, returns MD_OK (0)
 * or a negative MD_E* code, and never exits the process.  The message of the
 * last failure on the calling thread is md_last_error().  One engine handle is
 * not thread-safe; separate handles are independent.
 *
 * Units are the reference's: A, fs, kJ/mol, e, g/mol (src/simulation.cu:177,
 * src/forces.cu:47).  Boxes are orthorhombic and centred at the origin
 * (src/nve_leapfrog.cu:50-55).
 *
 * The entry points replace the single-GPU step path of the reference:
 *   md_set_forcefield   <- forces_init constant uploads      src/forces.cu:75-111
 *   md_set_topology     <- per-atom bond/angle lists          src/forces.cu:113-179
 *                          and the exclusion table            src/neighborlist.cu:64-99
 *   md_set_system /
 *   md_upload_state     <- config_init / config_HostToDevice  src/config.cu:22-40,56-62
 *   md_download_state   <- config_finish / _DeviceToHost      src/config.cu:42-77
 *   md_compute_forces   <- forces() + energy()                src/forces.cu:189-231
 *   md_run              <- SimulationRun_CUDA time loop       src/runGPU.cu:69-104
 *                          (neighborlist -> forces -> integrate, src/neighborlist.cu:109-146,
 *                           src/integrator.cu:131-221)
 *   md_get_neighbors    <- nnlist / nnlist_max readback        src/neighborlist.cu:111-115
 *   md_set_integrator   <- integrate_init ensemble/dt/T0      src/integrator.cu:47-97
 * The force-matching and batching entry points have no counterpart in the
 * reference (they extend the same step path).
 */
#ifndef MDGPU_MD_API_H
#define MDGPU_MD_API_H

#ifdef __cplusplus
extern "C" {
#endif

#define MD_OK 0
#define MD_E_INVALID (-1)      /* bad argument */
#define MD_E_HIP (-2)          /* HIP runtime error */
#define MD_E_STATE (-3)        /* call out of order (e.g. run before upload) */
#define MD_E_CAPACITY (-4)     /* a size limit was exceeded */
#define MD_E_UNSTABLE (-5)     /* non-finite coordinates or an atom left the box */

#define MD_MAX_TYPES 8          /* NUM_PARTICLE_TYPES, src/simulation.h:47 */
#define MD_MAX_BOND_TYPES 8     /* NUM_BOND_TYPES, src/simulation.h:48 */

/* Coulomb kinds = the reference's lkspace (src/simulation.cu:82-95) + none */
enum { MD_COUL_RF = 0, MD_COUL_WOLF = 1, MD_COUL_DAMPED = 2, MD_COUL_NONE = 3 };
enum { MD_LEAPFROG = 0, MD_VVERLET = 1 };
enum { MD_TH_NONE = 0, MD_TH_BERENDSEN = 1, MD_TH_NOSEHOOVER = 2 };
enum { MD_BARO_NONE = 0, MD_BARO_BERENDSEN = 1 };

typedef struct {
    int   ntype;                         /* 1..MD_MAX_TYPES */
    float lj1[MD_MAX_TYPES * MD_MAX_TYPES];   /* [ti*ntype+tj] = 48 eps sig^12, src/simulation.cu:191 */
    float lj2[MD_MAX_TYPES * MD_MAX_TYPES];   /* [ti*ntype+tj] = 24 eps sig^6,  src/simulation.cu:194 */
    int   coulomb;                       /* MD_COUL_* */
    float rc;                            /* cutoff, A */
    float alpha;                         /* Wolf / damped alpha, 1/A */
    float eps_rf;                        /* reaction-field dielectric */
    float skin;                          /* Verlet skin, A */
    float qfac;                          /* 1389.35427611722 (0 selects that default) */
    int   nbond_types;
    float bond_k2[MD_MAX_BOND_TYPES];    /* bnd1 = 2k (E = k (r-r0)^2), src/simulation.cu:227 */
    float bond_r0[MD_MAX_BOND_TYPES];
    int   nangle_types;
    float angle_k2[MD_MAX_BOND_TYPES];   /* ang1 = 2k, src/simulation.cu:242 */
    float angle_t0[MD_MAX_BOND_TYPES];   /* radians, src/simulation.cu:243 */
} md_forcefield;

typedef struct {
    int   integrator;   /* MD_LEAPFROG (vel = v(t-dt/2)) or MD_VVERLET (vel = v(t)) */
    int   thermostat;   /* MD_TH_*; leapfrog supports NONE and BERENDSEN */
    float dt;           /* fs */
    float tau;          /* fs; Berendsen time constant (reference tco = 0.01/fs -> 100 fs) or NH period */
    /* Pressure coupling (leapfrog only; the reference's NPT ensemble, enstype 2,
     * src/simulation.cu:71-75, whose kernel src/npt_leapfrog_ber.cu:28-123 never
     * scales the box): MD_BARO_BERENDSEN scales positions and box by
     * mu = cbrt(1 - compress dt / tau_p (P0 - P)) after every step, P = (sum m v^2
     * + W) / (3 V) from the step's kinetic energy and virial W (pairs and bonds).
     * Energies and virial are then computed at every step. */
    int   barostat;     /* MD_BARO_NONE (0) or MD_BARO_BERENDSEN */
    float press0;       /* P0, bar */
    float tau_p;        /* fs */
    float compress;     /* isothermal compressibility, 1/bar */
} md_integrator;

/* Per-step, per-replica statistics written by md_run (FP64 throughout). */
typedef struct {
    double ke;          /* kinetic energy, kJ/mol */
    double temp;        /* K, N_f = 3N-3, k_B = 8.314e-3 (src/integrator.cu:155) */
    double pe_lj;       /* LJ energy (shifted to 0 at rc), valid when has_energy */
    double pe_coul;     /* Coulomb pair energy, valid when has_energy */
    double pe_bond;
    double pe_angle;
    double virial;      /* sum over pairs and bonds of r.F, valid when has_energy */
    double conserved;   /* ke + pe (+ Nose-Hoover terms), valid when has_energy */
    double xi, vxi;     /* Nose-Hoover thermostat position / velocity */
    double volume;      /* box volume of the step's forces, A^3 */
    int    has_energy;
    int    pad;
} md_step_stats;

typedef struct md_engine md_engine;

const char* md_last_error(void);
/* library build string (arch, compiler) */
const char* md_version(void);

int md_create(md_engine** out, int device);
int md_destroy(md_engine* e);

int md_set_forcefield(md_engine* e, const md_forcefield* ff);
int md_set_integrator(md_engine* e, const md_integrator* in);

/* nrep independent replicas sharing one force field.  Atom arrays are the
 * replicas' atoms concatenated; natom[r], box[3r..3r+2], T0[r] per replica.
 * type: 0-based, charge: e, mass: g/mol per atom. */
int md_set_system(md_engine* e, int nrep, const int* natom, const float* box, const float* T0,
                  const int* type, const float* charge, const float* mass);
/* bonds: nb x (i, j, type); angles: na x (i, center j, k, type); global 0-based
 * atom indices.  1-2 bonded pairs are excluded from pair forces (<= 4 per atom). */
int md_set_topology(md_engine* e, int nb, const int* bonds, int na, const int* angles);
/* Rigid 3-site molecules (water), held by SETTLE (Miyamoto and Kollman,
 * J. Comput. Chem. 13, 952 (1992)) after every leapfrog step.  The reference's
 * constraint code (src/settle.c, src/settle.cu:1-178, src/xshake.cu:15-82) is
 * uncompiled pseudo-code; this is the published method.  mol: nmol x (O, H1, H2)
 * global 0-based atom indices (H1 and H2 of equal mass, all three in one
 * replica), held at the O-H distance d_oh and the H-H distance d_hh.  The atoms
 * of a molecule exclude each other's pair terms (with the 1-2 exclusions of
 * md_set_topology, at most 4 per atom); give no bonds or angles among them.
 * The temperature then counts 3N - 3 - 3 N_mol degrees of freedom per replica.
 * Runs need the leapfrog integrator (NVE or Berendsen) and no pressure coupling.
 * nmol = 0 removes the constraints.  Call after md_set_system and
 * md_set_topology (which keeps them). */
int md_set_rigid(md_engine* e, int nmol, const int* mol, float d_oh, float d_hh);

/* Smooth particle-mesh Ewald (Essmann et al., J. Chem. Phys. 103, 8577 (1995))
 * for the reciprocal-space part of the damped Coulomb term (MD_COUL_DAMPED, the
 * reference's lkspace 2, src/forces.cu:97-110, src/simulation.cu:90-94), at the
 * splitting parameter beta = the force field's alpha.  The reference reads kmax
 * for it and computes no reciprocal part; its GROMACS PME (src/pme.c) is vendored,
 * not built.  With PME on, each step adds the reciprocal forces (order-p
 * B-splines on a grid[0] x grid[1] x grid[2] mesh, FFTs by hipFFT), removes
 * -qfac qi qj erf(beta r)/r for every excluded pair, and adds the self and
 * neutralizing-background terms to the Coulomb energy and the virial.
 * order: 4, 6 or 8 (0 turns PME off); grid: points per box edge, >= order, at
 * most 1024 (sizes with factors 2, 3, 5 transform fastest).  Runs need one
 * replica, charged atoms and MD_COUL_DAMPED (checked at the next md_run /
 * md_compute_forces); under pressure coupling the long-range part follows the
 * box of each step.  Force matching has no PME term. */
int md_set_pme(md_engine* e, const int* grid, int order);

/* pos (N x 3, required), vel (N x 3 or NULL = zero).  Builds the neighbor lists. */
int md_upload_state(md_engine* e, const float* pos, const float* vel);
/* Any pointer may be NULL.  Arrays are N x 3 in the uploaded atom order. */
int md_download_state(md_engine* e, float* pos, float* vel, float* force);

/* Forces at the current positions (N x 3, may be NULL); energy per replica
 * nrep x 4 (lj, coul, bond, angle) and virial per replica (may be NULL). */
int md_compute_forces(md_engine* e, float* force, double* energy, double* virial);

/* Advance nsteps.  energy_every > 0 computes energies at steps with
 * (step index % energy_every == 0); stats (nsteps x nrep, may be NULL).
 * Work is queued as hipGraph segments; the call returns when done. */
int md_run(md_engine* e, int nsteps, int energy_every, md_step_stats* stats);

/* Full neighbor lists in uploaded atom indices: counts[N], nbrs[N x cap]
 * (unsorted within a row).  Returns MD_E_CAPACITY if cap is too small
 * (counts still complete). */
int md_get_neighbors(md_engine* e, int cap, int* counts, int* nbrs);

/* Consistency of the device's sorted atom order and neighbor lists against the
 * positions of the last list build (test support; no counterpart in the
 * reference, whose list has no guard, src/neighborlist.cu:284).  counts[MD_NCHECK]
 * receives the number of violations of each invariant, all zero when sound:
 *   [0] an atom slot s with slot_of[ids[s]] != s (the order is not a bijection)
 *   [1] a row entry outside its tile's neighbor set
 *   [2] a listed pair at distance >= r_list at the build positions, or i with itself
 *   [3] a tile-set entry that is not an atom of the tile's replica
 *   [4] a row or tile set beyond its capacity
 *   [5] row padding that is not the empty entry */
#define MD_NCHECK 6
int md_check_lists(md_engine* e, int* counts);

/* Sums over all atoms of the current forces and of the momentum m v, FP64
 * (the force and momentum lines of the reference's statistics print,
 * src/prints.cu:117-118), reduced on the device: 2 x 3 doubles come back.
 * Either pointer may be NULL. */
int md_get_sums(md_engine* e, double* force_sum, double* momentum_sum);

/* Current box of every replica (3 floats each): the box of md_set_system, or
 * after a run with pressure coupling the coupled box (src/simulation.h:54
 * config.box, which the reference would have had to update for NPT,
 * src/integrator.cu:60-61). */
int md_get_box(md_engine* e, float* box);

/* Canonical FP32 thresholds the kernels use (for the parity checker):
 * rc2 = rc*rc, rl2 = (float)((skin + sqrt(rc2))^2), skin2q = skin*skin/4. */
int md_get_thresholds(md_engine* e, float* rc2, float* rl2, float* skin2q);

/* Engine counters: neighbor-list rebuilds per replica summed, steps run,
 * list capacity per atom, current lanes-per-atom and grid of the step kernel. */
int md_get_info(md_engine* e, long long* rebuilds, long long* steps, int* cap, int* lanes, int* grid);

/* Neighbor-list capacity (entries per atom), the reference's nmaxlist
 * (.inp, src/neighborlist.cu:61); at most 65535.  The lists are rebuilt at the next call
 * that needs them; a row that does not fit does not corrupt anything (the
 * reference has no guard): the run that met it is rolled back to its start
 * and replayed with a larger capacity. */
int md_set_list_capacity(md_engine* e, int cap);

/* Launch configuration (no counterpart in the reference, whose launch shape
 * is fixed by src/gpu_properties.cu:31-67).  Setting it rebuilds the lists
 * when the cell grid changes. */
typedef struct md_launch_config {
    int lanes;         /* lanes cooperating on one atom in the step kernel: 4, 8, 16 (a workgroup of
                          64 x lanes threads per 64-atom tile) */
    int grid;          /* unused (0): the step kernel runs one workgroup per 64-atom tile */
    int cell_div;      /* cell edge >= r_list / cell_div, 1..3 */
    int check_every;   /* steps queued between two (device-gated) list-rebuild chains */
    int build_lanes;   /* wavefronts per 64-atom tile in the list build: 4 (256 threads, more
                          workgroups resident per CU) or 8 (512 threads); the same lists */
    int submit;        /* 0: replayed HIP graph segments, 1: direct kernel launches */
    int refresh;       /* 1: also rebuild a list after check_every steps, ahead of the skin trip */
    int build_kernel;  /* 0: the tile neighbor-set build (the only list build) */
    int defer;         /* leapfrog: 1 = each step kernel finalizes the previous step of its block
                          (temperature, thermostat, statistics) instead of a finalize kernel per step */
    int vv_fuse;       /* NVE velocity Verlet: 1 = one step kernel per step ends the step and begins the
                          next (leapfrog form, so defer applies); 0 = kick, force and finalize kernels */
    int fm_fused;      /* force matching without bonded terms: 1 = one pass computes the force and the
                          gradient sums; 0 = a force pass, then the adjoint pass */
    int persist;       /* small single systems (one replica of at most 4,096 atoms, Lennard-Jones only,
                          no bonded terms, leapfrog or NVE velocity Verlet, no pressure coupling):
                          1 = every step of a run in one launch of a few workgroups that meet at a
                          grid barrier each step (1024 threads each, `lanes` per atom); other systems
                          ignore it */
} md_launch_config;
int md_set_launch(md_engine* e, const md_launch_config* cfg);
int md_get_launch(md_engine* e, md_launch_config* cfg);
/* Time candidate launch configurations on the current system and keep the
 * fastest.  nsteps is the length of the runs the caller will make (md_run
 * calls): each trial times a sequence of runs of that length (at least 384
 * steps in all), so a caller of short runs gets the configuration that is
 * fastest for short runs.  The physical state is restored after every trial.
 * best_ms_per_step may be NULL. */
int md_autotune(md_engine* e, int nsteps, double* best_ms_per_step);

/* The persistent kernel (md_launch_config.persist; no counterpart in the
 * reference): runs made on it, replays of a run after a lane's row outgrew its
 * length (the run restarts from its first step with rows twice as long), and
 * the current row length per lane. */
int md_persist_info(md_engine* e, long long* runs, long long* replays, int* row_len);

/* Kernel timing with HIP events recorded on the engine's stream around every
 * launch of the pair/step kernel during md_run (the persistent kernel, which
 * runs all steps of a run in one launch, counts as one launch per step).
 * enable resets the counters. */
int md_timing_enable(md_engine* e, int enable);
int md_timing_get(md_engine* e, double* step_kernel_ms, long long* step_kernel_launches,
                  double* rebuild_ms, long long* rebuild_count);
/* Force `reps` full list rebuilds of every replica at the current positions
 * and report the mean time of one (k_force_trip + the five rebuild kernels),
 * from HIP events on the engine stream.  No counterpart in the reference. */
int md_rebuild_bench(md_engine* e, int reps, double* ms_per_rebuild);
/* Plain device copy of nbytes (calibration of HBM counters). Returns ms. */
int md_copy_bench(md_engine* e, long long nbytes, int reps, double* ms);

/* ---- force matching (no counterpart in the reference) ----
 * Frames share the engine's system (type, charge, box of replica 0).
 * md_fm_set_frames uploads nframes x N x 3 positions and target forces once. */
int md_fm_set_frames(md_engine* e, int nframes, const float* pos, const float* target_force);
/* loss = mean over frames*atoms*3 of (F - F*)^2 with charges scaled by qscale
 * and LJ tables lj1/lj2 (ntype x ntype, ordered).  Gradients per ordered table
 * entry and for qscale, FP64.  Any gradient pointer may be NULL. */
int md_fm_loss_grad(md_engine* e, const float* lj1, const float* lj2, float qscale,
                    double* loss, double* g_lj1, double* g_lj2, double* g_qscale);

/* ---- no-allocation variant (no counterpart in the reference) ----
 * For callers that cannot allocate device memory at evaluation time (code
 * under HIP graph capture, tight fitting loops, hosts that forbid hipMalloc):
 * the same evaluation and results as md_fm_loss_grad, but the per-tile
 * partial sums, the trial LJ table and the final sums live in one fixed
 * MD_FM_SCRATCH_BYTES device array of the library (one copy per GPU, held
 * until the process ends, its address resolved once per device by
 * md_create).  Nothing is allocated on the device or by a library routine.
 * Returns MD_E_CAPACITY when ntiles x (2 + 2 k^2) x 8 bytes (k = 1, 2 or 4 for
 * 1, 2 or 3-4 atom types; ntiles = the frames' 64-atom tiles) exceed the array
 * less its first 2 KiB: 13,081 tiles for two types, e.g. 204 frames of 4,096
 * ions.  Every call on one GPU shares that array and leaves its contents
 * behind (each call writes every slot it reads): calls that overlap on one
 * GPU, from several engines or threads, race. */
#define MD_FM_SCRATCH_BYTES (1 << 20)
int md_fm_loss_grad_noalloc(md_engine* e, const float* lj1, const float* lj2, float qscale,
                            double* loss, double* g_lj1, double* g_lj2, double* g_qscale);

#ifdef __cplusplus
}
#endif
#endif
