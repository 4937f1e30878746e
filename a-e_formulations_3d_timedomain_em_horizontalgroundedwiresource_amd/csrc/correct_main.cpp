// Drop-in executable for CorrectTimeMesh.exe: corrects too-short decades of infite.0 in the
// current directory (or the one given with -d).
#include <cstdio>
#include <cstring>

#include "../../include/emtd.h"

int main(int argc, char** argv) {
    const char* dir = ".";
    for (int i = 1; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], "-d")) dir = argv[++i];
    const int rc = emtd_correct_time_mesh(dir);
    if (rc) std::fprintf(stderr, "CorrectTimeMesh: error 0x%04x: %s\n", rc, emtd_last_error());
    return rc;
}
