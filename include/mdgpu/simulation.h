/*
 * simulation.h -- the reference's host-level run API (C++ linkage), backed
 * by the MI355X engine (md_api.h).  A program written against
 * src/simulation.h:132-138 and src/runGPU.h:33 of md_gpu links against
 * libmdgpu.so instead.
 *
 * The data model re-declares the reference structs (src/simulation.h:50-130)
 * field for field with plain POD vector types (md_float4 / md_uint2 /
 * md_uint4 carry the same x,y,z,w members and the same 16 / 8 / 16-byte
 * alignment as CUDA's float4 / uint2 / uint4), so a caller that allocates the
 * structs itself, as src/simulation.cu:40-131 does, passes them unchanged.
 * Units and conversions are the reference's: mass x 1e4 (src/simulation.cu:177),
 * lj1 = 48 eps sig^12 / lj2 = 24 eps sig^6 (:191,194), bnd1 = 2k (:227),
 * ang1 = 2k and ang2 in radians (:242-243).
 *
 * Behaviour kept: file names <name>.inp/.fld/.cfg in, <name>.xyz and
 * <name>.xyz.cfg out (src/simulation.cu:323,331); positional parsing with
 * ignored labels; time loop t = 0..timrun inclusive (src/runGPU.cu:69); a
 * statistics line "t\tT\tKE" every timprn steps (src/prints.cu:39) followed by
 * the force / momentum sums (src/prints.cu:117-118).
 * Behaviour changed on purpose: errors are reported with a message and
 * SimulationInit returns NULL (the reference read unchecked files); NPT
 * (enstype 2, read but never integrated by the reference: its kernel
 * src/npt_leapfrog_ber.cu:28-123 is only called from commented-out code,
 * src/integrator.cu:157-164) runs leapfrog with Berendsen temperature and
 * pressure coupling to press_0 (bar), and the coupled box is written back to
 * the config.
 *
 * Optional trailing .inp lines (after "rdel:"), ignored by older files:
 *   integrator: leapfrog | vverlet
 *   thermostat: none | berendsen | nosehoover  <tau_fs>
 *   barostat: berendsen <tau_p_fs> <compressibility_per_bar>   (NPT; defaults 1000 fs, 4.5e-5)
 * SimulationInit keeps these choices in library-side state tied to the integr
 * it returns (the reference's integr has no fields for them, and its tcons is
 * never set, src/integrator.cu:68).  An integr the library did not create runs
 * the reference's own scheme: leapfrog, NVE for enstype 0, Berendsen with the
 * reference's tco = 0.01/fs (tau 100 fs, src/integrator.cu:146) for enstype 1.
 */
#ifndef MDGPU_SIMULATION_H
#define MDGPU_SIMULATION_H

#define LINE_LENGTH 80
#define NUM_PARTICLE_TYPES 8
#define NUM_BOND_TYPES 8

typedef struct __attribute__((aligned(16))) { float x, y, z, w; } md_float4;
typedef struct __attribute__((aligned(8))) { unsigned int x, y; } md_uint2;
typedef struct __attribute__((aligned(16))) { unsigned int x, y, z, w; } md_uint4;

typedef struct {            /* src/simulation.h:51-59 */
    int natom;
    int* itype;
    md_float4 box;
    md_float4 boxi;
    md_float4* pos;         /* x, y, z, charge */
    md_float4* vel;         /* vx, vy, vz, mass x 1e4 */
    md_float4* force;
} config;

typedef struct {            /* src/simulation.h:62-86 */
    int ntype;
    int* nt;
    char** name;
    float* mass;
    float* charge;
    float* lj1;
    float* lj2;
    float* bnd1;
    float* bnd2;
    float* ang1;
    float* ang2;
    int tbonds;
    int tangle;
    int nrigid;
    int nbonds;
    int nangle;
    int nmaxbond;
    unsigned int* bndlist_max;
    md_uint2* bndlist;
    md_uint4* iatn;         /* (w = type, x = i, y = j), 1-based as read */
    unsigned int* anglist_max;
    md_uint4* anglist;
    md_uint4* iang;         /* (w = type, x = i, y = center j, z = k), 1-based */
} ffield;

typedef struct {            /* src/simulation.h:88-100 */
    int timeq;
    int timrun;
    int timprn;
    float rcutsq;
    float reacf;
    int lkspace;
    int kmax;
    float alfa;
    float skin;
    int nmaxlist;
    float rdel;
} params;

typedef struct {            /* src/simulation.h:102-106 */
    md_float4 tdm;
    md_float4 energ_nonbond;
    md_float4 energ_bonded;
} thermodyn;

typedef struct {            /* src/simulation.h:108-122 */
    char* ensemble;
    int enstype;            /* 0 NVE, 1 NVT, 2 NPT */
    float dt;
    float temper_0;
    float tcons;            /* not read by the run (the reference never sets it) */
    float press_0;
    float pcons;
    float temper;
    float* Ke_sum;
    float* Wi_sum;
    float* Ke_total;
    float* Wi_total;
    thermodyn* therm;
} integr;

typedef struct {            /* src/simulation.h:124-130 */
    char* simname;
    params* pars;
    integr* integ;
    config* conf;
    ffield* fld;
} simulation;

simulation* SimulationInit(char* name_sim);
void SimulationRun(simulation* sim);
void SimulationFinish(simulation* sim);
ffield* FFieldInput(char* name);
void FFieldFinish(ffield* fld);
config* ConfigInput(char* name, float* mass);
void ConfigFinish(config* cfg, ffield* fld, char* name);

/* GPU driver with the reference signature (src/runGPU.h:33); returns 0 or an
 * md_api.h error code (the reference returned void and exited on error). */
int SimulationRun_CUDA(params* pars, integr* integ, ffield* fld, config* conf);

/* Status of the last SimulationInit / SimulationRun on this process (0 = ok);
 * the athom tool maps a failure to exit code 1. */
int SimulationLastStatus(void);

#endif
