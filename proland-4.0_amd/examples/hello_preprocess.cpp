// A C++ caller of the drop-in, written the way Proland's apps call the
// reference at init (src/atmo/examples/atmo/HelloWorld.cpp:240,
// src/demo/ProlandDemo.cpp:404): include PreprocessAtmo.h, pass default
// AtmoParameters and an output directory.
//
//   hello_preprocess <output dir>   writes transmittance.raw, irradiance.raw and
//                                   inscatter.raw there (or keeps them when
//                                   inscatter.raw exists); exit status = the
//                                   preprocessAtmoStatus() code
//   hello_preprocess --defaults     writes the 96 bytes of AtmoParameters() to stdout
#include <cstdio>
#include <cstring>

#include "PreprocessAtmo.h"
#include "atmo.h"

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <output dir> | --defaults\n", argv[0]);
        return 2;
    }
    if (std::strcmp(argv[1], "--defaults") == 0) {
        const proland::AtmoParameters p;
        return std::fwrite(&p, sizeof p, 1, stdout) == 1 ? 0 : 1;
    }
    proland::preprocessAtmo(proland::AtmoParameters(), argv[1]);
    const int st = proland::preprocessAtmoStatus();
    if (st != 0) std::fprintf(stderr, "preprocessAtmo: status %d: %s\n", st, atmo_last_error());
    return st;
}
