// Drop-in executable for SummatorEDMN.exe: "summatoredmn WithDDV profile place [-d DIR]"
// (M/SummatorEDMN/main.cpp:247-295).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/emtd.h"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "Error: need 3 parameters\n1 - flag WithDDV\n2 - number of profile\n"
                             "3 - number of place by profile\n");
        return 1;
    }
    const char* dir = ".";
    for (int i = 4; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], "-d")) dir = argv[++i];
    const int rc = emtd_summator_edmn(dir, std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]));
    if (rc) std::fprintf(stderr, "SummatorEDMN: error 0x%04x: %s\n", rc, emtd_last_error());
    return rc;
}
