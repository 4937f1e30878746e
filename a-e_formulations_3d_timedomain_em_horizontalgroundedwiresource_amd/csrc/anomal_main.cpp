// Drop-in executable for UnloadAnomalNonstat.exe: the anomalous-edge export of the calculation
// directory (the current one, or the one given with -d).
#include <cstdio>
#include <cstring>

#include "../../include/emtd.h"

int main(int argc, char** argv) {
    const char* dir = ".";
    for (int i = 1; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], "-d")) dir = argv[++i];
    const int rc = emtd_unload_anomal_nonstat(dir);
    if (rc) std::fprintf(stderr, "UnloadAnomalNonstat: error 0x%04x: %s\n", rc, emtd_last_error());
    return rc;
}
