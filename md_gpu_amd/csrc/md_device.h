// md_device.h -- data layout and kernel argument records shared by the HIP
// kernels and the host engine.
//
// Layout in HBM (all arrays sorted along the space-filling curve; "slot" =
// position in that order):
//   pos[2][S]   float4 (x, y, z, q)   double-buffered: a step reads one, writes the other.
//               Coordinates are wrapped into the box at every list build and run
//               free in between (an atom is at most skin/2 outside), so the image
//               shifts the build recorded stay valid until the next build.
//   vel[S]      float4 (vx, vy, vz, 1/m_int)
//   force[S]    float4 (fx, fy, fz, 0)
//   ids[S]      int    original atom index (-1 = padding slot)
//   typ[S]      int    atom type
// Neighbor list, per 64-slot tile t (a tile never straddles two replicas):
//   jset[t][k]  uint32 the tile's neighbor set: slot | image shift code << 24 | type << 29,
//               k < jcount[t]; local index k + 1 (local index 0 is a far-away dummy)
//   nl16[S][cap] uint16 per-atom rows: local index into the tile's set times 16 (the
//               byte offset of the staged entry; sets hold at most MD_JMAX entries), in
//               the block layout of lpos16(); entries past the row length up to the end
//               of the row's last block are 0 (the dummy), ncount[S] = row length
//   rperm[S]    uint8 per tile: the tile's atoms (0..63) by decreasing row length; the
//               step kernel gives lane group g the atom rperm[g], so the atoms of a
//               wavefront have rows of similar length
// The step kernel stages the tile's set into LDS once per step, each position
// already shifted to its image, so a pair needs no minimum-image arithmetic.
// Replicas occupy 64-aligned slot ranges, so a wavefront never straddles two.
//
// Control: every replica carries a RepState on the device.  Kernels decide
// per replica whether they run (list valid, phase, step target), so a graph
// segment can be replayed without host decisions; the host only checks the
// step counters when the queued work should be done.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MD_TILE 64
#define MD_JBITS 29             // bond_list entries: partner | type << 29
#define MD_JMASK 0x1FFFFFFFu
#define MD_SLOTBITS 24          // neighbor-set entries: slot | shift code << 24 | type << 29
#define MD_SLOTMASK 0xFFFFFFu
#define MD_SHIFT0 13            // shift code of the unshifted image: (kx+1) + 3 (ky+1) + 9 (kz+1)
#define MD_KB_REF 8.314e-3      // src/integrator.cu:155 (1000/8.314)
#define MD_BAR_PER_KJMOL_A3 16605.390671738467   // 1 kJ/mol/A^3 in bar
#define MD_NSTAT 8              // accumulators per replica
#define MD_CTAB 256             // intervals of the Coulomb force table on [0, rc]
#define MD_SUBSHIFT 25          // list build: ranks and original indices stay below 2^25 (the slot limit is
                                // lower); the eighth of the cell an atom sits in is packed above them
#define MD_FAR 1.0e5f           // coordinate of the dummy neighbor (local index 0)
#define MD_JMAX 4095            // largest tile neighbor set (a row entry is the local index times 16)

enum { ACC_KE2 = 0, ACC_LJ, ACC_COUL, ACC_BOND, ACC_ANGLE, ACC_VIR };
enum { MODE_FORCE = 0, MODE_LF = 1, MODE_VV = 2 };

struct RepParam {
    float L[3], Linv[3], halfL[3];
    float cinv[3];     // nc / L
    float cedge[3];    // L / nc
    int nc[3];
    int cell_off;      // offset of this replica's cells in the cell tables
    int cs_off;        // offset in cell_start (= cell_off + replica index)
    int slot_off, nslot, natom, atom_off;
    float T0;
    float nf;          // 3N - 3, less 3 per rigid molecule
    double kT0, Q;     // Nose-Hoover: kB T0, N_f kB T0 tau^2
    float Lb[3];       // box at the replica's last list build (pressure coupling scales L)
    float pad_;
};

// Per-replica control state (device).
struct RepState {
    long long step;    // steps completed (absolute)
    long long target;  // step count at which the current run stops
    long long start;   // first step of the current run (statistics index base)
    int phase;         // velocity Verlet: 0 = between steps, 1 = kicked/drifted, forces pending
    int nb;            // completed neighbor-list builds
    int trip[2];       // trip[nb & 1] != 0: the list is stale for the current positions
    int age;           // steps completed since the last build
    int ran;           // deferred-finalize copies: the step kernel ran this step for the replica
    double lam;        // deferred-finalize copies: Berendsen factor for the replica's next step
};

// Device-wide counters that let a launch with nothing to do exit on its first
// load.  They only gate early exits: every kernel still checks each replica.
struct Gate {
    int n_stale;       // replicas whose list is stale
    int n_run;         // leapfrog: replicas that can step; velocity Verlet: replicas that can kick
    int n_force;       // velocity Verlet: replicas whose force half is pending (list valid)
    int pad;
};

// Storage position of logical entry e in a row of 16-bit entries read by G lanes
// per atom.  Rows are stored in blocks of 8G entries; inside a block lane l's
// 16-byte load holds logical entries l, l+G, ..., l+7G, so each LDS read of the
// atom's G lanes covers G consecutive entries (mostly consecutive staged
// neighbors: a row lists the members of one cell together).
__host__ __device__ inline int lpos16(int e, int G) {
    return (e & ~(8 * G - 1)) + 8 * (e % G) + ((e / G) & 7);
}
__device__ __forceinline__ uint32_t lpos16_u(uint32_t e, int gs) {
    const uint32_t G = 1u << gs;
    return (e & ~(8u * G - 1u)) | ((e & (G - 1u)) << 3) | __builtin_amdgcn_ubfe(e, (uint32_t)gs, 3u);
}

// image shift (kx, ky, kz) in {-1, 0, 1}^3 of a shift code
__host__ __device__ inline int shift_code(int kx, int ky, int kz) { return (kx + 1) + 3 * (ky + 1) + 9 * (kz + 1); }

// h at table coordinate x = r * MD_CTAB / rc >= 0 from the cubic Coulomb table in
// LDS (MD_CTAB intervals and an end entry h(rc); beyond rc the value is finite
// and unused).  x - (int)x is exact for x >= 0, so fract gives the offset.
__device__ __forceinline__ float ctab_eval(const float4* s_ct, float x) {
    const int k = min((int)x, MD_CTAB);
    const float t = __builtin_amdgcn_fractf(x);
    const float4 c4 = s_ct[k];
    return fmaf(fmaf(fmaf(c4.w, t, c4.z), t, c4.y), t, c4.x);
}

// Compiler-only ordering point (no instruction is emitted): the loads above it
// are not moved below it, e.g. into the branch after an early exit that is the
// only user of their values, so that they are in flight together.
__device__ __forceinline__ void loads_issued() { __asm__ __volatile__("" ::: "memory"); }

// Workgroups are dispatched round-robin over MI355X's 8 XCDs (workgroup b on
// XCD b mod 8), each with its own L2.  xcd_block maps the launch index b to a
// logical index so that every XCD takes one contiguous range of the n logical
// workgroups (a bijection on [0, n)): neighboring atoms, and the positions their
// rows gather, stay in one L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int x = b & 7, k = b >> 3, q = n >> 3, rem = n & 7;
    return x * q + min(x, rem) + k;
}

// current position buffer of a replica
__host__ __device__ inline int cur_buf(const RepState& s) { return (int)((s.step + s.phase) & 1); }

// the list of the current epoch is stale (a select, not an indexed load, so a
// register copy of the state never has to live in scratch)
__host__ __device__ inline bool stale(const RepState& s) { return ((s.nb & 1) ? s.trip[1] : s.trip[0]) != 0; }

// Position wrapped into the centred box, the reference's x -= L trunc(2x/L)
// (src/nve_leapfrog.cu:50-55); valid for |x| < 1.5 L.
__host__ __device__ inline float wrap_coord(float x, float L, float Linv) {
    return fmaf(-L, truncf(x * 2.0f * Linv), x);
}

// Thermostat state per replica (FP64).  S is the velocity scale the next
// position-writing kernel applies; *_end keep the state after the end-of-step
// half so a run can stop there.
struct ThState {
    double xi, vxi;
    double S;          // pending scale (LF: lambda for next step; VV: s_end * s_begin)
    double S_end;      // VV: end-of-step half scale only
    double xi_end, vxi_end;
    double ke2_cur;    // VV: 2 KE of the velocities after the end-of-step half
    double pad;
};

// Step statistics record (mirrors md_step_stats in md_api.h)
struct StatRec {
    double ke, temp, pe_lj, pe_coul, pe_bond, pe_angle, virial, conserved, xi, vxi, volume;
    int has_energy, pad;
};

struct StepArgs {
    float4* pos[2];
    float4* vel;
    float4* force;
    const int* ids;
    const int* typ;
    const uint16_t* nl16;     // [S][cap] rows of local indices x 16 (lpos16 layout)
    const int* ncount;
    const uint8_t* rperm;     // [S] per tile: atoms by decreasing row length
    const uint32_t* jset;     // [ntiles][jcap] tile neighbor sets
    const int* jcount;        // [ntiles]
    const float4* pos_build;
    const int* tile_rep;
    RepParam* rep;        // the box changes under pressure coupling (k_step_finalize)
    RepState* st;
    Gate* gate;
    ThState* th;
    double* part;         // [2][MD_NSTAT][ntiles] FP64 partial sums, one slot per tile
                          // (buffer par; the deferred finalize reads buffer par ^ 1)
    RepState* lst;        // [2][R] state handed from one step launch to the next (deferred finalize)
    int vvfused;          // NVE velocity Verlet in leapfrog form (mode MODE_LF): the step kernel ends the
                          // step (second half-kick) and, unless the run ends, begins the next one (first
                          // half-kick and drift); between launches the replica is kicked (phase 1)
    int lagmode;          // leapfrog block with deferred finalize: the step kernel writes lst[par]
    int open;             // the block's rebuild chain had no commit: the first step kernel (lag = 0)
                          // opens the new list epoch of every stale (just rebuilt) replica
    int lag;              // the previous launch was a step kernel of the same block: finalize it first
    int prefilter;        // many replicas and no deferred finalize: each tile first reads its replica's
                          // state and leaves if the replica does not step in this launch
    int par;              // buffer parity of this launch
    int* err;             // [1] non-finite coordinates
    StatRec* stats;       // [stat_cap][R] or null
    const float4* ljtab;  // [8*8] (lj1, lj2, LJ energy shift, 0)
    const float4* ctab;   // [MD_CTAB + 1] cubic coefficients of h(r) = A(r) - r A(rc)/rc (Wolf / damped)
    const int* slot_of;   // [N] original index -> slot
    const uint32_t* bond_list;  // [N*4] (partner | type << 29)
    const int* nbond;           // [N]
    const int4* ang_list;       // [N*4] (role | type << 8, a, b, 0)
    const int* nang;            // [N]
    const float4* target;       // [S] target forces in slot order (force matching)
    int cap;              // row capacity in entries (a multiple of 128)
    int jcap;             // neighbor-set capacity per tile
    int ntype;
    int nrep;
    int ntiles;           // S / 64
    int mode;             // MODE_*
    int thermostat;
    int bonded;
    int energy_every;     // energies at steps with step % energy_every == 0 (0: never)
    int force_energy;     // MODE_FORCE: compute energies
    int stat_cap;         // records per replica in stats
    float rc2, qfac, palfa, spiialfa, alfasq, wolf, reacf;
    float ecut;           // Coulomb energy shift: Wolf E(rc)/rc, RF c_rf, else 0
    float log2rc;         // log2(rc) for the Wolf energy's ln(r/rc) term
    float ct_scale;       // MD_CTAB / rc: table index per unit r
    float qq_scale;       // multiplies q_i q_j (force matching)
    float dt, tau;
    float skin2q;
    int max_age;          // > 0: the list is also refreshed after this many steps
    float bk2[8], br0[8], ak2[8], at0[8];
    // Berendsen pressure coupling (leapfrog, finalize kernel): baro != 0 scales
    // positions and box by mu = cbrt(1 - baro_k (P0 - P)) after each step
    int baro;
    float p0;             // bar
    float baro_k;         // compress dt / tau_p, 1/bar
    float rl, rc;         // list radius and cutoff (the skin check under a scaled box)
    // smooth PME (damped Coulomb kernel only): the reciprocal and excluded-pair force
    // of each slot, added to the pair force, and the step's energy and virial of the
    // long-range part, added to the replica's first tile; null without PME
    const float4* frec;
    const double* pme_e;
};

// Smooth particle-mesh Ewald for one replica (pme.hip), a chain of kernels and
// two FFTs before each step kernel: spread the charges onto the grid (fixed
// point, so the sum does not depend on the order of the atomic adds), convert,
// transform, multiply by the influence function (energy and virial on the
// way), transform back, gather the forces.
#define MD_PME_MAX_ORDER 8
#define MD_PME_QUNIT 4294967296.0   // fixed-point units per charge unit in the spread grid
struct PmeArgs {
    const float4* pos[2];       // current buffer pos[cur_buf(st[0])]
    const int* ids;             // [S] original index or -1
    const int* slot_of;         // [N]
    const int4* excl;           // [N] excluded partners (-1 empty)
    const RepState* st;         // the replica (one)
    unsigned long long* qacc;   // [K0 K1 K2] charge grid in fixed point, zero between steps
    float* qgrid;               // [K0 K1 K2] charges (transform input); dE/dQ (inverse output)
    float2* sgrid;              // [K0 K1 (K2/2 + 1)] the transform
    const float2* tab;          // [K0 + K1 + K2] per axis and index: (signed frequency m, |b(m)|^2)
    float4* frec;               // [S] reciprocal + excluded-pair force per slot
    double* part;               // [nconv + ngather][2] energy and virial per workgroup
    double* out;                // [2] energy and virial of the long-range part
    int K0, K1, K2;
    int nslot;
    int mode;                   // MODE_*: which replicas step (the step kernel's gate)
    int has_excl;
    int nconv, ngather;         // workgroups of the influence and gather kernels
    int order;                  // B-spline order (4, 6 or 8)
    const RepParam* rep;        // the replica's box, read every step (pressure coupling changes it)
    double ik0, ik1, ik2;       // 1 / K
    float beta;
    float qfac;
    float pi2_b2;               // pi^2 / beta^2
    double e_self;              // -qfac beta / sqrt(pi) sum q^2
    double qsum2;               // (sum q)^2: the neutralizing-background term -qfac pi qsum2 / (2 V beta^2)
};

// Progress of a run for the host: one small copy per check instead of one per
// array (tmax holds the builds' longest rows in [0, ntiles) and the largest
// tile sets in [ntiles, 2 ntiles), since the host last cleared it).
struct RunReport {
    long long min_step;   // smallest step counter over the replicas
    int rowmax;           // longest row of any build
    int setmax;           // largest tile neighbor set of any build
    int err;              // non-finite coordinates or an atom beyond the box
    int pad;
};

// Streaming kernel of velocity Verlet (begin of step): v = S v + F dt/2m; x += v dt
struct KickArgs {
    float4* pos[2];
    float4* vel;
    const float4* force;
    const int* ids;
    const float4* pos_build;
    const int* tile_rep;
    const RepParam* rep;
    RepState* st;
    Gate* gate;
    const ThState* th;
    int* err;
    int* tflag;           // [S/64] skin check tripped in the 64-slot tile
    int nslot;
    int nrep;
    float dt, skin2q;
    int apply_drift;      // 0: run end (v *= S_end only)
    int lfgate;           // the step kernel is in leapfrog form: a kicked replica stays counted in n_run
};

// SETTLE for rigid 3-site molecules (settle.hip), between a leapfrog step kernel
// and its finalize.
struct SettleArgs {
    float4* pos[2];
    float4* vel;
    const float4* pos_build;
    const int* ids;
    const int* tile_rep;
    const RepParam* rep;
    const RepState* st;
    const int* slot_of;
    const int* rig_of;    // [N] molecule index of an O atom, -2 for an H atom, -1 unconstrained
    const int4* rig_mol;  // [nmol] (O, H1, H2, 0) original atom indices
    double* part;         // the step kernel's partial slots (KE and skin-trip planes rewritten)
    int* err;
    int par;              // parity of the step launch's slots
    int ntiles;
    float dt, skin2q;
    float rc_hh;          // half the H-H distance
    float h;              // height of the triangle: sqrt(d_OH^2 - rc_hh^2)
};

// Persistent kernel for small single systems (persist.hip): all steps of a run
// in one launch of at most MD_PK_WG workgroups of MD_PK_THREADS threads.
#define MD_PK_SLOTS 4096          // largest system (slots): the staged positions fill at most 64 KB of LDS
#define MD_PK_WG 64               // most workgroups (one wavefront polls every barrier flag)
#define MD_PK_THREADS 1024
#define MD_PK_TIMEOUT 50000000ull // barrier wait limit, 100 MHz wall-clock ticks (0.5 s)
#define MD_PK_ERR_ROWS 4          // err bit: a lane's row outgrew cap_l (the run is replayed with longer rows)
#define MD_PK_ERR_WAIT 8          // err bit: a barrier wait timed out

struct PersistArgs {
    float4* pos[2];
    float4* vel;
    float4* force;
    const int* ids;
    const int* typ;
    const float4* ljtab;  // [8*8] (lj1, lj2, LJ energy shift, 0)
    RepState* st;         // replica 0
    ThState* th;
    const RepParam* rep;
    unsigned* flags;      // [2][MD_PK_WG] barrier flags, zero at launch
    double* pub;          // [2][MD_PK_WG][4] partial sums published at the barriers (2 KE, LJ, virial)
    int* err;             // 1: an atom left the box; MD_PK_ERR_*
    StatRec* stats;       // [stat_cap] or null
    int stat_cap;
    int nslot;
    int nwg;              // ceil(nslot / (MD_PK_THREADS / lanes))
    long long nsteps;
    int energy_every;
    int leapfrog;         // 1: leapfrog (thermostat none or Berendsen); 0: NVE velocity Verlet
    int thermostat;
    int ntype;
    int cap_l;            // row entries per lane
    float dt, tau, rc2, rl2, skin2q;
};

struct BuildArgs {
    float4* pos[2];       // current buffer per replica: pos[cur_buf(st[r])]
    float4* vel;
    float4* force;
    int* ids;
    int* typ;
    RepParam* repw;       // the same parameters, written: Lb at the build
    // scratch, new order
    float4* pos_s;
    float4* vel_s;
    float4* force_s;
    int* ids_s;
    int* typ_s;
    float4* pos_build;
    // cells
    int* key;             // [S] global cell index of each (old) slot
    int* rank;            // [S] arrival rank in its cell
    int* tmp;             // [S] new slot -> old slot (cell order, arrival order)
    int* tmpid;           // [S] new slot -> original index
    int* cell_count;      // [C]
    int* cell_start;      // [C + R]
    const int* cell_rank; // [C] linear cell -> curve rank (per replica block)
    int* slot_of;         // [N]
    uint16_t* nl16;
    int* ncount;
    uint8_t* rperm;
    uint32_t* jset;
    int* jcount;
    int* tilemax;         // [S/64] longest row per 64-slot tile over the builds since the host cleared it
    int* jmax;            // [S/64] largest neighbor set per tile over the same builds
    const int4* excl;     // [N] 1-2 partners (-1 empty)
    RepState* st;
    Gate* gate;
    const int* tile_rep;
    const RepParam* rep;
    int nslot;
    int nrep;
    int cap;
    int jcap;
    float rl2;
    int has_excl;
    int row_g;            // lanes per atom of the step kernel (row layout, lpos16)
    int build_nw;         // wavefronts per tile in the tile build (4 or 8; the same lists)
    int lfgate;           // the step kernel is in leapfrog form: n_run counts every valid replica short of
                          // its target, whatever its phase
};
