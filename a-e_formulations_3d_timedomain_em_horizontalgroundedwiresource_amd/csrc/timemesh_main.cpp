// Drop-in executable for TimeDirectSolver.exe: regularizes infite.0 and resamples the source
// tables in the current directory (or the one given with -d).
#include <cstdio>
#include <cstring>

#include "../../include/emtd.h"

int main(int argc, char** argv) {
    const char* dir = ".";
    for (int i = 1; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], "-d")) dir = argv[++i];
    const int rc = emtd_time_direct_solver(dir);
    if (rc) std::fprintf(stderr, "TimeDirectSolver: error 0x%04x: %s\n", rc, emtd_last_error());
    return rc;
}
