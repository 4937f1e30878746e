// Drop-in executable for CalcNonstat3DLine.exe: runs the module in the current directory
// (or the directory given with -d). Variant flags: -A/-E (formulation), -2/-3 (scheme);
// without them the file "emtd_variant" decides. -fp32 selects FP32 factor + FP64 refinement.
#include <cstdio>
#include <cstring>

#include "../../include/emtd.h"

int main(int argc, char** argv) {
    const char* dir = ".";
    int form = -1, scheme = -1, prec = EMTD_FP64, dev = -1;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-A")) form = EMTD_FORM_A;
        else if (!std::strcmp(argv[i], "-E")) form = EMTD_FORM_E;
        else if (!std::strcmp(argv[i], "-2")) scheme = 2;
        else if (!std::strcmp(argv[i], "-3")) scheme = 3;
        else if (!std::strcmp(argv[i], "-fp32")) prec = EMTD_FP32_IR;
        else if (!std::strcmp(argv[i], "-d") && i + 1 < argc) dir = argv[++i];
        else if (!std::strcmp(argv[i], "-device") && i + 1 < argc) dev = std::atoi(argv[++i]);
    }
    int rc = emtd_calc_nonstat_3d_line(dir, form, scheme, prec, dev);
    if (rc) std::fprintf(stderr, "CalcNonstat3DLine: error 0x%04x: %s\n", rc, emtd_last_error());
    return rc;
}
