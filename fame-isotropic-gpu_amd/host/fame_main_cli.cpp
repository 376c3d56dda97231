/*
 * fame_main_cli.cpp -- the reference's Main.cpp:8-49 as a program: reads
 * Popt.txt and FAME_Material_Data/<name>.txt from the directory given as the
 * first argument (default: the working directory), runs the solver through
 * FAME_Main_Code and writes Data_Omega_array.txt and Data_Profile.txt there.
 */
#include "fame_driver.h"

#include <cstdio>

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : ".";
    const int rc = fame_run(dir);
    if (rc != 0) std::fprintf(stderr, "fame_main: failed (%d)\n", rc);
    return rc == 0 ? 0 : 1;
}
