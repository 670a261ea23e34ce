// athom.cpp -- command-line tool with the reference's interface
// (src/athom.cu:30-47): `athom <name>` reads <name>.inp/.fld/.cfg, runs the
// MD on the GPU and writes <name>.xyz and <name>.xyz.cfg.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mdgpu/simulation.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <name>   (reads <name>.inp, <name>.fld, <name>.cfg)\n", argv[0]);
        return 1;
    }
    std::vector<char> name(strlen(argv[1]) + LINE_LENGTH);
    strcpy(name.data(), argv[1]);
    simulation* sim = SimulationInit(name.data());
    if (!sim) return 1;
    SimulationRun(sim);
    int rc = SimulationLastStatus();
    SimulationFinish(sim);
    return rc == 0 ? 0 : 1;
}
