// ref_layout_caller.cpp -- test program: a caller of SimulationRun_CUDA that
// knows only the reference's host data model.  It declares the structs with
// the reference's layout (src/simulation.h:51-122, CUDA's float4 / uint2 /
// uint4 being 16 / 8 / 16-byte aligned), allocates and fills them the way
// src/simulation.cu:40-131 and :144-316 do (integr from malloc: tcons, Ke_sum
// and the other buffers hold whatever the heap had), links against
// libmdgpu.so and calls the reference's entry point (src/runGPU.h:33).
//
//   ref_layout_caller <in.txt> <out.txt>
// in.txt (written by tests/test_gpu_cli.py):
//   natom ntype
//   box_x box_y box_z
//   ensemble temper_0 dt timrun timprn
//   rc lkspace alfa skin nmaxlist
//   mass[ntype] (g/mol), lj1[ntype*ntype], lj2[ntype*ntype]
//   natom lines: type(1-based) charge x y z vx vy vz
// out.txt: natom lines x y z vx vy vz (after the run)
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>

struct __attribute__((aligned(16))) float4 { float x, y, z, w; };
struct __attribute__((aligned(8))) uint2 { unsigned int x, y; };
struct __attribute__((aligned(16))) uint4 { unsigned int x, y, z, w; };

typedef struct {
    int natom;
    int* itype;
    float4 box;
    float4 boxi;
    float4* pos;
    float4* vel;
    float4* force;
} config;

typedef struct {
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
    uint2* bndlist;
    uint4* iatn;
    unsigned int* anglist_max;
    uint4* anglist;
    uint4* iang;
} ffield;

typedef struct {
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

typedef struct {
    float4 tdm;
    float4 energ_nonbond;
    float4 energ_bonded;
} thermodyn;

typedef struct {
    char* ensemble;
    int enstype;
    float dt;
    float temper_0;
    float tcons;
    float press_0;
    float pcons;
    float temper;
    float* Ke_sum;
    float* Wi_sum;
    float* Ke_total;
    float* Wi_total;
    thermodyn* therm;
} integr;

void SimulationRun_CUDA(params* pars, integr* integ, ffield* fld, config* conf);

static_assert(sizeof(float4) == 16 && alignof(float4) == 16, "CUDA float4");
static_assert(offsetof(config, box) == 16 && offsetof(config, pos) == 48, "config layout");

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: ref_layout_caller in.txt out.txt\n");
        return 2;
    }
    FILE* f = fopen(argv[1], "r");
    if (!f) return 2;
    int natom, ntype;
    char ens[16];
    params* pars = (params*)malloc(sizeof(params));
    integr* integ = (integr*)malloc(sizeof(integr));
    memset(integ, 0xA5, sizeof(integr));   // the heap's leftovers: tcons, Ke_sum, ... are garbage
    config* conf = (config*)malloc(sizeof(config));
    ffield* fld = (ffield*)malloc(sizeof(ffield));
    memset(fld, 0, sizeof(ffield));
    float rc;
    if (fscanf(f, "%d %d", &natom, &ntype) != 2) return 2;
    if (fscanf(f, "%f %f %f", &conf->box.x, &conf->box.y, &conf->box.z) != 3) return 2;
    if (fscanf(f, "%15s %f %f %d %d", ens, &integ->temper_0, &integ->dt, &pars->timrun, &pars->timprn) != 5) return 2;
    if (fscanf(f, "%f %d %f %f %d", &rc, &pars->lkspace, &pars->alfa, &pars->skin, &pars->nmaxlist) != 5) return 2;
    integ->ensemble = (char*)malloc(16);
    strcpy(integ->ensemble, ens);
    integ->enstype = strcmp(ens, "nvt") == 0 ? 1 : 0;                  // src/simulation.cu:62-75
    pars->timeq = 0;
    pars->rcutsq = rc * rc;                                            // :97
    pars->reacf = 0.0f;
    pars->rdel = 0.1f;
    conf->natom = natom;
    conf->boxi.x = 1.0f / conf->box.x;                                 // :290-292
    conf->boxi.y = 1.0f / conf->box.y;
    conf->boxi.z = 1.0f / conf->box.z;
    fld->ntype = ntype;
    fld->mass = (float*)malloc(8 * sizeof(float));
    fld->lj1 = (float*)malloc(64 * sizeof(float));
    fld->lj2 = (float*)malloc(64 * sizeof(float));
    fld->bnd1 = (float*)calloc(8, sizeof(float));
    fld->bnd2 = (float*)calloc(8, sizeof(float));
    fld->ang1 = (float*)calloc(8, sizeof(float));
    fld->ang2 = (float*)calloc(8, sizeof(float));
    for (int t = 0; t < ntype; t++) {
        if (fscanf(f, "%f", &fld->mass[t]) != 1) return 2;
        fld->mass[t] *= 10000.0f;                                      // :177
    }
    for (int k = 0; k < ntype * ntype; k++)
        if (fscanf(f, "%f", &fld->lj1[k]) != 1) return 2;
    for (int k = 0; k < ntype * ntype; k++)
        if (fscanf(f, "%f", &fld->lj2[k]) != 1) return 2;
    conf->itype = (int*)malloc(natom * sizeof(int));
    conf->pos = (float4*)malloc(natom * sizeof(float4));
    conf->vel = (float4*)malloc(natom * sizeof(float4));
    conf->force = (float4*)malloc(natom * sizeof(float4));
    for (int i = 0; i < natom; i++) {                                  // :305-312
        int t;
        if (fscanf(f, "%d %f %f %f %f %f %f %f", &t, &conf->pos[i].w, &conf->pos[i].x, &conf->pos[i].y,
                   &conf->pos[i].z, &conf->vel[i].x, &conf->vel[i].y, &conf->vel[i].z) != 8)
            return 2;
        conf->itype[i] = t - 1;
        conf->vel[i].w = fld->mass[t - 1];
    }
    fclose(f);
    SimulationRun_CUDA(pars, integ, fld, conf);
    FILE* o = fopen(argv[2], "w");
    if (!o) return 2;
    for (int i = 0; i < natom; i++)
        fprintf(o, "%.9g %.9g %.9g %.9g %.9g %.9g\n", conf->pos[i].x, conf->pos[i].y, conf->pos[i].z, conf->vel[i].x,
                conf->vel[i].y, conf->vel[i].z);
    fclose(o);
    return 0;
}
