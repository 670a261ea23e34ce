// simulation.cpp -- the reference's SimulationInit / SimulationRun /
// SimulationFinish API and file formats (src/simulation.cu:40-345,
// src/runGPU.cu:43-114, src/prints.cu:37-119), implemented over the C ABI of
// the MI355X engine.  Host code only.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mdgpu/md_api.h"
#include "mdgpu/simulation.h"

namespace {

int g_status = MD_OK;
const float kPiRef = 3.1415927f;   // M_Pi of src/simulation.h:37

// Integrator choice of an integr made by SimulationInit (the optional .inp keys).
// The reference's integr has no fields for it, so it is kept here, keyed by the
// struct's address, and dropped by SimulationFinish.
struct RunOptions {
    int integrator = MD_LEAPFROG;   // the reference loop runs leapfrog (vver = 0, src/runGPU.cu:71)
    int thermostat = -1;            // -1: the ensemble's default
    float tau = 100.0f;             // reference tco = 0.01 /fs (src/integrator.cu:146)
    float tau_p = 1000.0f;          // NPT: Berendsen pressure coupling time, fs
    float compress = 4.5e-5f;       // NPT: isothermal compressibility, 1/bar (water)
    int pme_K[3] = {0, 0, 0};       // "pme: K0 K1 K2 order" (lkspace 2 only): smooth PME grid
    int pme_order = 0;              // and B-spline order; 0 = the reference's real-space term only
};
std::mutex g_opt_mu;
std::map<const integr*, RunOptions> g_opts;

RunOptions options_of(const integr* in, bool* found) {
    std::lock_guard<std::mutex> lk(g_opt_mu);
    auto it = g_opts.find(in);
    *found = it != g_opts.end();
    return *found ? it->second : RunOptions{};
}

bool rd(int got, int want, const char* what, const std::string& file) {
    if (got != want) {
        fprintf(stderr, "md_gpu: %s: cannot read %s\n", file.c_str(), what);
        return false;
    }
    return true;
}

template <class T>
T* alloc(size_t n) {
    return static_cast<T*>(calloc(n ? n : 1, sizeof(T)));
}

}  // namespace

int SimulationLastStatus(void) { return g_status; }

// ---------------------------------------------------------------- .fld
ffield* FFieldInput(char* name) {
    std::string file = std::string(name) + ".fld";
    FILE* fin = fopen(file.c_str(), "r");
    if (!fin) {
        fprintf(stderr, "md_gpu: cannot open %s\n", file.c_str());
        return nullptr;
    }
    char saux[256];
    ffield* fld = alloc<ffield>(1);
    bool ok = rd(fscanf(fin, "%255s %d", saux, &fld->ntype), 2, "number of atom types", file);
    if (ok && (fld->ntype < 1 || fld->ntype > NUM_PARTICLE_TYPES)) {
        fprintf(stderr, "md_gpu: %s: %d atom types (1..%d supported)\n", file.c_str(), fld->ntype, NUM_PARTICLE_TYPES);
        ok = false;
    }
    const int nty = NUM_PARTICLE_TYPES;
    fld->nt = alloc<int>(nty);
    fld->name = alloc<char*>(nty);
    fld->mass = alloc<float>(nty);
    fld->charge = alloc<float>(nty);
    fld->lj1 = alloc<float>(nty * nty);
    fld->lj2 = alloc<float>(nty * nty);
    for (int it = 0; ok && it < fld->ntype; it++) {
        int iaux;
        fld->name[it] = alloc<char>(LINE_LENGTH);
        ok = rd(fscanf(fin, "%d %d", &iaux, &fld->nt[it]), 2, "atom type header", file) &&
             rd(fscanf(fin, "%79s %f %f", fld->name[it], &fld->mass[it], &fld->charge[it]), 3, "atom type", file);
        fld->mass[it] *= 10000.0;   // src/simulation.cu:177
    }
    int nvdw = 0;
    if (ok) ok = rd(fscanf(fin, "%255s %d", saux, &nvdw), 2, "VDW count", file);
    for (int k = 0; ok && k < nvdw; k++) {   // src/simulation.cu:184-205
        int i, j;
        float a, b;
        char form[32];
        ok = rd(fscanf(fin, "%d %d %31s %f %f", &i, &j, form, &a, &b), 5, "VDW line", file);
        if (!ok) break;
        i--;
        j--;
        if (i < 0 || j < 0 || i >= fld->ntype || j >= fld->ntype) {
            fprintf(stderr, "md_gpu: %s: VDW pair %d %d outside the atom types\n", file.c_str(), i + 1, j + 1);
            ok = false;
            break;
        }
        float l1, l2;
        if (strcmp(form, "LJ") == 0) {          // sigma, epsilon
            l1 = 12.0f * 4.0f * b * pow(a, 12);
            l2 = 6.0f * 4.0f * b * pow(a, 6);
        } else if (strcmp(form, "12-6") == 0) { // A, B
            l1 = 12.0f * a;
            l2 = 6.0f * b;
        } else {
            fprintf(stderr, "md_gpu: %s: unknown VDW form %s\n", file.c_str(), form);
            ok = false;
            break;
        }
        fld->lj1[i * fld->ntype + j] = fld->lj1[j * fld->ntype + i] = l1;
        fld->lj2[i * fld->ntype + j] = fld->lj2[j * fld->ntype + i] = l2;
    }
    fld->bnd1 = alloc<float>(NUM_BOND_TYPES);
    fld->bnd2 = alloc<float>(NUM_BOND_TYPES);
    fld->ang1 = alloc<float>(NUM_BOND_TYPES);
    fld->ang2 = alloc<float>(NUM_BOND_TYPES);
    if (ok) ok = rd(fscanf(fin, "%255s %d", saux, &fld->tbonds), 2, "bond type count", file);
    if (ok && (fld->tbonds < 0 || fld->tbonds > NUM_BOND_TYPES)) ok = false;
    for (int it = 0; ok && it < fld->tbonds; it++) {
        int iaux;
        ok = rd(fscanf(fin, "%d %f %f", &iaux, &fld->bnd1[it], &fld->bnd2[it]), 3, "bond type", file);
        fld->bnd1[it] *= 2.0f;   // src/simulation.cu:227
    }
    if (ok) ok = rd(fscanf(fin, "%255s %d", saux, &fld->tangle), 2, "angle type count", file);
    if (ok && (fld->tangle < 0 || fld->tangle > NUM_BOND_TYPES)) ok = false;
    for (int it = 0; ok && it < fld->tangle; it++) {
        int iaux;
        ok = rd(fscanf(fin, "%d %f %f", &iaux, &fld->ang1[it], &fld->ang2[it]), 3, "angle type", file);
        fld->ang1[it] *= 2.0f;              // src/simulation.cu:242
        fld->ang2[it] *= kPiRef / 180.0f;   // :243
    }
    if (ok) ok = rd(fscanf(fin, "%255s %d", saux, &fld->nbonds), 2, "bond count", file);
    if (ok) fld->iatn = alloc<md_uint4>(fld->nbonds);
    for (int it = 0; ok && it < fld->nbonds; it++) {
        int iaux;
        ok = rd(fscanf(fin, "%d %u %u %u", &iaux, &fld->iatn[it].w, &fld->iatn[it].x, &fld->iatn[it].y), 4, "bond", file);
    }
    if (ok) ok = rd(fscanf(fin, "%255s %d", saux, &fld->nangle), 2, "angle count", file);
    if (ok) fld->iang = alloc<md_uint4>(fld->nangle);
    for (int it = 0; ok && it < fld->nangle; it++) {
        int iaux;
        ok = rd(fscanf(fin, "%d %u %u %u %u", &iaux, &fld->iang[it].w, &fld->iang[it].x, &fld->iang[it].y,
                       &fld->iang[it].z), 5, "angle", file);
    }
    fclose(fin);
    if (!ok) {
        FFieldFinish(fld);
        return nullptr;
    }
    fld->nmaxbond = 4;
    return fld;
}

void FFieldFinish(ffield* fld) {
    if (!fld) return;
    if (fld->name)
        for (int i = 0; i < NUM_PARTICLE_TYPES; i++) free(fld->name[i]);
    free(fld->name); free(fld->nt); free(fld->mass); free(fld->charge); free(fld->lj1); free(fld->lj2);
    free(fld->bnd1); free(fld->bnd2); free(fld->ang1); free(fld->ang2); free(fld->iatn); free(fld->iang);
    free(fld);
}

// ---------------------------------------------------------------- .cfg
config* ConfigInput(char* name, float* mass) {
    std::string file = std::string(name) + ".cfg";
    FILE* fin = fopen(file.c_str(), "r");
    if (!fin) {
        fprintf(stderr, "md_gpu: cannot open %s\n", file.c_str());
        return nullptr;
    }
    config* conf = alloc<config>(1);
    int cftyp = 0;
    bool ok = rd(fscanf(fin, "%d %d", &conf->natom, &cftyp), 2, "atom count", file) &&
              rd(fscanf(fin, "%f %f %f", &conf->box.x, &conf->box.y, &conf->box.z), 3, "box", file);
    if (ok && conf->natom < 1) ok = false;
    if (ok) {
        conf->boxi.x = 1.0 / conf->box.x;   // src/simulation.cu:290-292
        conf->boxi.y = 1.0 / conf->box.y;
        conf->boxi.z = 1.0 / conf->box.z;
        conf->itype = alloc<int>(conf->natom);
        conf->pos = alloc<md_float4>(conf->natom);
        conf->vel = alloc<md_float4>(conf->natom);
        conf->force = alloc<md_float4>(conf->natom);
    }
    for (int i = 0; ok && i < conf->natom; i++) {   // src/simulation.cu:305-312
        int iaux;
        ok = rd(fscanf(fin, "%d %d %f", &iaux, &conf->itype[i], &conf->pos[i].w), 3, "atom header", file);
        if (!ok) break;
        conf->itype[i] -= 1;
        if (conf->itype[i] < 0 || conf->itype[i] >= NUM_PARTICLE_TYPES) {
            fprintf(stderr, "md_gpu: %s: atom %d has type %d\n", file.c_str(), i + 1, conf->itype[i] + 1);
            ok = false;
            break;
        }
        conf->vel[i].w = mass ? mass[conf->itype[i]] : 0.0f;
        ok = rd(fscanf(fin, "%f %f %f", &conf->pos[i].x, &conf->pos[i].y, &conf->pos[i].z), 3, "position", file);
        if (ok && cftyp > 0)
            ok = rd(fscanf(fin, "%f %f %f", &conf->vel[i].x, &conf->vel[i].y, &conf->vel[i].z), 3, "velocity", file);
    }
    fclose(fin);
    if (!ok) {
        free(conf->itype); free(conf->pos); free(conf->vel); free(conf->force); free(conf);
        return nullptr;
    }
    return conf;
}

void ConfigFinish(config* conf, ffield* fld, char* name) {
    if (!conf) return;
    std::string xyz = std::string(name) + ".xyz";        // src/simulation.cu:323
    FILE* fout = fopen(xyz.c_str(), "w");
    if (fout) {
        fprintf(fout, "%d\n", conf->natom);
        fprintf(fout, "%f %f %f\n", conf->box.x, conf->box.y, conf->box.z);
        for (int i = 0; i < conf->natom; i++)
            fprintf(fout, "%s %f %f %f\n", fld && fld->name && fld->name[conf->itype[i]] ? fld->name[conf->itype[i]] : "X",
                    conf->pos[i].x, conf->pos[i].y, conf->pos[i].z);
        fclose(fout);
    } else {
        fprintf(stderr, "md_gpu: cannot write %s\n", xyz.c_str());
    }
    std::string cfg = xyz + ".cfg";                      // the reference's strcat chain, :331
    fout = fopen(cfg.c_str(), "w");
    if (fout) {
        fprintf(fout, "%d %d\n", conf->natom, 1);
        fprintf(fout, "%f %f %f\n", conf->box.x, conf->box.y, conf->box.z);
        for (int i = 0; i < conf->natom; i++) {
            fprintf(fout, "%d %d %f\n", i + 1, conf->itype[i] + 1, conf->pos[i].w);
            fprintf(fout, "%f %f %f\n", conf->pos[i].x, conf->pos[i].y, conf->pos[i].z);
            fprintf(fout, "%f %f %f\n", conf->vel[i].x, conf->vel[i].y, conf->vel[i].z);
        }
        fclose(fout);
    } else {
        fprintf(stderr, "md_gpu: cannot write %s\n", cfg.c_str());
    }
    free(conf->itype); free(conf->pos); free(conf->vel); free(conf->force); free(conf);
}

// ---------------------------------------------------------------- .inp
simulation* SimulationInit(char* name_sim) {
    g_status = MD_OK;
    std::string file = std::string(name_sim) + ".inp";
    FILE* fin = fopen(file.c_str(), "r");
    if (!fin) {
        fprintf(stderr, "md_gpu: cannot open %s\n", file.c_str());
        g_status = MD_E_INVALID;
        return nullptr;
    }
    simulation* sim = alloc<simulation>(1);
    sim->simname = alloc<char>(strlen(name_sim) + 1);
    strcpy(sim->simname, name_sim);
    params* pars = alloc<params>(1);
    integr* integ = alloc<integr>(1);
    RunOptions opt;
    char saux[256];
    integ->ensemble = alloc<char>(16);
    integ->tcons = 100.0f;
    bool ok = rd(fscanf(fin, "%255s %15s", saux, integ->ensemble), 2, "ensemble", file);
    if (ok && strcmp(integ->ensemble, "nvt") == 0) {
        integ->enstype = 1;
        ok = rd(fscanf(fin, "%255s %f", saux, &integ->temper_0), 2, "temperature", file);
    } else if (ok && strcmp(integ->ensemble, "npt") == 0) {
        integ->enstype = 2;
        ok = rd(fscanf(fin, "%255s %f %f", saux, &integ->temper_0, &integ->press_0), 3, "temperature/pressure", file);
    }
    if (ok) ok = rd(fscanf(fin, "%255s %d %d %d %f", saux, &pars->timeq, &pars->timrun, &pars->timprn, &integ->dt), 5,
                    "time steps", file);
    if (ok) ok = rd(fscanf(fin, "%255s %f %d", saux, &pars->rcutsq, &pars->lkspace), 3, "cutoff", file);
    if (ok) {
        if (pars->lkspace == 0) {        // reaction field, src/simulation.cu:82-87
            float eps;
            ok = rd(fscanf(fin, "%255s %f", saux, &eps), 2, "reaction-field dielectric", file);
            pars->reacf = 2.0f * (eps - 1.0f) / (2.0f * eps + 1.0f) / pow(pars->rcutsq, 3);
        } else if (pars->lkspace == 1) { // Wolf
            ok = rd(fscanf(fin, "%255s %f", saux, &pars->alfa), 2, "alpha", file);
        } else {                         // damped real-space Ewald term
            ok = rd(fscanf(fin, "%255s %d", saux, &pars->kmax), 2, "kmax", file);
            pars->alfa = 3.09 / pars->rcutsq;
        }
    }
    if (ok) pars->rcutsq *= pars->rcutsq;
    if (ok) ok = rd(fscanf(fin, "%255s %f %d", saux, &pars->skin, &pars->nmaxlist), 3, "skin", file);
    if (ok) ok = rd(fscanf(fin, "%255s %f", saux, &pars->rdel), 2, "rdel", file);
    // optional trailing keys
    while (ok && fscanf(fin, "%255s", saux) == 1) {
        char val[64];
        if (strcmp(saux, "integrator:") == 0 && fscanf(fin, "%63s", val) == 1) {
            if (strcmp(val, "vverlet") == 0) opt.integrator = MD_VVERLET;
            else if (strcmp(val, "leapfrog") == 0) opt.integrator = MD_LEAPFROG;
            else { fprintf(stderr, "md_gpu: %s: unknown integrator %s\n", file.c_str(), val); ok = false; }
        } else if (strcmp(saux, "thermostat:") == 0 && fscanf(fin, "%63s %f", val, &opt.tau) == 2) {
            integ->tcons = opt.tau;
            if (strcmp(val, "none") == 0) opt.thermostat = MD_TH_NONE;
            else if (strcmp(val, "berendsen") == 0) opt.thermostat = MD_TH_BERENDSEN;
            else if (strcmp(val, "nosehoover") == 0) opt.thermostat = MD_TH_NOSEHOOVER;
            else { fprintf(stderr, "md_gpu: %s: unknown thermostat %s\n", file.c_str(), val); ok = false; }
        } else if (strcmp(saux, "barostat:") == 0 && fscanf(fin, "%63s %f %f", val, &opt.tau_p, &opt.compress) == 3) {
            if (strcmp(val, "berendsen") != 0) {
                fprintf(stderr, "md_gpu: %s: unknown barostat %s\n", file.c_str(), val);
                ok = false;
            } else if (!(opt.tau_p > 0) || !(opt.compress >= 0)) {
                fprintf(stderr, "md_gpu: %s: barostat needs tau_p > 0 and compressibility >= 0\n", file.c_str());
                ok = false;
            }
        } else if (strcmp(saux, "pme:") == 0 &&
                   fscanf(fin, "%d %d %d %d", &opt.pme_K[0], &opt.pme_K[1], &opt.pme_K[2], &opt.pme_order) == 4) {
            if (pars->lkspace != 2) {
                fprintf(stderr, "md_gpu: %s: pme needs the Ewald real-space term (lkspace 2)\n", file.c_str());
                ok = false;
            }
        } else {
            fprintf(stderr, "md_gpu: %s: unknown key %s\n", file.c_str(), saux);
            ok = false;
        }
    }
    fclose(fin);
    if (ok) {
        sim->fld = FFieldInput(name_sim);
        ok = sim->fld != nullptr;
    }
    if (ok) {
        sim->conf = ConfigInput(name_sim, sim->fld->mass);
        ok = sim->conf != nullptr;
    }
    sim->pars = pars;
    sim->integ = integ;
    if (!ok) {
        g_status = MD_E_INVALID;
        FFieldFinish(sim->fld);
        free(pars); free(integ->ensemble); free(integ); free(sim->simname); free(sim);
        return nullptr;
    }
    integ->temper = integ->temper_0;
    integ->Ke_sum = alloc<float>(1);
    {
        std::lock_guard<std::mutex> lk(g_opt_mu);
        g_opts[integ] = opt;
    }
    return sim;
}

// ---------------------------------------------------------------- run
int SimulationRun_CUDA(params* pars, integr* integ, ffield* fld, config* conf) {
    auto die = [&](md_engine* e, int rc) {
        fprintf(stderr, "md_gpu error: %s\n", md_last_error());
        if (e) md_destroy(e);
        g_status = rc;
        return rc;
    };
    md_engine* e = nullptr;
    int rc = md_create(&e, 0);
    if (rc) return die(nullptr, rc);
    printf("%s\n", md_version());
    md_forcefield ff;
    memset(&ff, 0, sizeof ff);
    const int nt = fld->ntype;
    ff.ntype = nt;
    for (int i = 0; i < nt * nt; i++) {
        ff.lj1[i] = fld->lj1[i];
        ff.lj2[i] = fld->lj2[i];
    }
    const float rc_ = std::sqrt(pars->rcutsq);
    ff.coulomb = pars->lkspace == 0 ? MD_COUL_RF : (pars->lkspace == 1 ? MD_COUL_WOLF : MD_COUL_DAMPED);
    ff.rc = rc_;
    ff.alpha = pars->alfa;
    if (pars->lkspace == 0) {   // recover eps_rf from reacf = 2(eps-1)/((2eps+1) rc^3)
        double k = 0.5 * (double)pars->reacf * rc_ * rc_ * rc_;
        ff.eps_rf = (float)((1.0 + k) / (1.0 - 2.0 * k));
    }
    ff.skin = pars->skin;
    ff.qfac = 0.0f;
    ff.nbond_types = fld->tbonds;
    for (int i = 0; i < fld->tbonds; i++) { ff.bond_k2[i] = fld->bnd1[i]; ff.bond_r0[i] = fld->bnd2[i]; }
    ff.nangle_types = fld->tangle;
    for (int i = 0; i < fld->tangle; i++) { ff.angle_k2[i] = fld->ang1[i]; ff.angle_t0[i] = fld->ang2[i]; }
    if ((rc = md_set_forcefield(e, &ff))) return die(e, rc);
    // integrator and thermostat: the options SimulationInit recorded for this integr,
    // else the reference's scheme (tcons is not read: the reference never sets it)
    bool own = false;   // the integr was made by SimulationInit (its Ke_sum is a host buffer)
    const RunOptions opt = options_of(integ, &own);
    md_integrator in;
    in.integrator = opt.integrator;
    in.thermostat = integ->enstype == 0 ? MD_TH_NONE : (opt.thermostat >= 0 ? opt.thermostat : MD_TH_BERENDSEN);
    in.dt = integ->dt;
    in.tau = opt.tau > 0 ? opt.tau : 100.0f;
    // NPT (enstype 2, src/simulation.cu:71-75): Berendsen pressure coupling to press_0 (bar)
    in.barostat = integ->enstype == 2 ? MD_BARO_BERENDSEN : MD_BARO_NONE;
    in.press0 = integ->enstype == 2 ? integ->press_0 : 1.0f;
    in.tau_p = opt.tau_p;
    in.compress = opt.compress;
    if ((rc = md_set_integrator(e, &in))) return die(e, rc);
    const int N = conf->natom;
    std::vector<int> type(N);
    std::vector<float> q(N), m(N), pos(3 * (size_t)N), vel(3 * (size_t)N), frc(3 * (size_t)N);
    for (int i = 0; i < N; i++) {
        type[i] = conf->itype[i];
        q[i] = conf->pos[i].w;
        m[i] = conf->vel[i].w / 10000.0f;
        pos[3 * i] = conf->pos[i].x; pos[3 * i + 1] = conf->pos[i].y; pos[3 * i + 2] = conf->pos[i].z;
        vel[3 * i] = conf->vel[i].x; vel[3 * i + 1] = conf->vel[i].y; vel[3 * i + 2] = conf->vel[i].z;
    }
    float box[3] = {conf->box.x, conf->box.y, conf->box.z};
    float T0 = integ->temper_0;
    if ((rc = md_set_system(e, 1, &N, box, &T0, type.data(), q.data(), m.data()))) return die(e, rc);
    std::vector<int> bonds, angles;
    for (int b = 0; b < fld->nbonds; b++) {
        bonds.push_back((int)fld->iatn[b].x - 1); bonds.push_back((int)fld->iatn[b].y - 1);
        bonds.push_back((int)fld->iatn[b].w - 1);
    }
    for (int a = 0; a < fld->nangle; a++) {
        angles.push_back((int)fld->iang[a].x - 1); angles.push_back((int)fld->iang[a].y - 1);
        angles.push_back((int)fld->iang[a].z - 1); angles.push_back((int)fld->iang[a].w - 1);
    }
    if ((rc = md_set_topology(e, fld->nbonds, bonds.data(), fld->nangle, angles.data()))) return die(e, rc);
    if (opt.pme_order > 0 && (rc = md_set_pme(e, opt.pme_K, opt.pme_order))) return die(e, rc);
    if ((rc = md_upload_state(e, pos.data(), vel.data()))) return die(e, rc);
    // small single Lennard-Jones systems run every step of an md_run call in one
    // launch (the engine ignores the option for any other system)
    md_launch_config lc;
    if ((rc = md_get_launch(e, &lc))) return die(e, rc);
    lc.persist = 1;
    if ((rc = md_set_launch(e, &lc))) return die(e, rc);

    printf("Starting the main loop ...\n");
    const int prn = pars->timprn > 0 ? pars->timprn : 1;
    std::vector<md_step_stats> st;
    for (int t = 0; t <= pars->timrun;) {   // t = 0..timrun inclusive, src/runGPU.cu:69
        // run through the next printed step (the smallest multiple of prn >= t) or timrun
        const int next = ((t + prn - 1) / prn) * prn;
        const int k = std::min(next, pars->timrun) - t + 1;
        st.resize(k);
        if ((rc = md_run(e, k, prn, st.data()))) return die(e, rc);
        t += k;
        const int last = t - 1;
        if (last % prn == 0) {
            const md_step_stats& s = st[k - 1];
            integ->temper = (float)s.temp;
            const float ke = (float)s.ke;
            if (own) integ->Ke_sum[0] = ke;   // a caller's integr may hold no host buffer here
            printf("%i\t%f\t%f\n", last, integ->temper, ke);   // src/prints.cu:39
            double fs[3], ps[3];   // reduced on the device: no state download per print
            if ((rc = md_get_sums(e, fs, ps))) return die(e, rc);
            printf("force: %f %f %f\n", fs[0], fs[1], fs[2]);       // src/prints.cu:117-118
            printf("moment: %f %f %f\n", ps[0], ps[1], ps[2]);
            if (s.has_energy)
                printf("energy: %f %f %f %f %f\n", s.pe_lj, s.pe_coul, s.pe_bond, s.pe_angle, s.conserved);
        }
    }
    if ((rc = md_download_state(e, pos.data(), vel.data(), frc.data()))) return die(e, rc);
    if (in.barostat != MD_BARO_NONE) {   // the coupled box goes back to the caller's config
        float b[3];
        if ((rc = md_get_box(e, b))) return die(e, rc);
        conf->box.x = b[0]; conf->box.y = b[1]; conf->box.z = b[2];
        conf->boxi.x = 1.0f / b[0]; conf->boxi.y = 1.0f / b[1]; conf->boxi.z = 1.0f / b[2];
    }
    for (int i = 0; i < N; i++) {
        conf->pos[i].x = pos[3 * i]; conf->pos[i].y = pos[3 * i + 1]; conf->pos[i].z = pos[3 * i + 2];
        conf->vel[i].x = vel[3 * i]; conf->vel[i].y = vel[3 * i + 1]; conf->vel[i].z = vel[3 * i + 2];
        conf->force[i].x = frc[3 * i]; conf->force[i].y = frc[3 * i + 1]; conf->force[i].z = frc[3 * i + 2];
    }
    md_destroy(e);
    return MD_OK;
}

void SimulationRun(simulation* sim) {
    printf("About to run simulation on GPU ...\n");
    auto t0 = std::chrono::steady_clock::now();
    int rc = SimulationRun_CUDA(sim->pars, sim->integ, sim->fld, sim->conf);
    double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc == MD_OK) printf("Finished simulation.\n Total time: %f s.\n", el);
    fflush(stdout);
}

void SimulationFinish(simulation* sim) {
    if (!sim) return;
    // the reference frees the force field before ConfigFinish reads its type
    // names (src/simulation.cu:139-140); here the configuration is written first
    ConfigFinish(sim->conf, sim->fld, sim->simname);
    FFieldFinish(sim->fld);
    free(sim->pars);
    if (sim->integ) {
        {
            std::lock_guard<std::mutex> lk(g_opt_mu);
            g_opts.erase(sim->integ);
        }
        free(sim->integ->ensemble);
        free(sim->integ->Ke_sum);
        free(sim->integ);
    }
    free(sim->simname);
    free(sim);
}
