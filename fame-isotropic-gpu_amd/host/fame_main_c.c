/*
 * fame_main_c.c -- the C-linkage symbol FAME_Main_Code, for C and FFI callers
 * (ctypes, cgo, JNI).  C++ callers bind the C++-linkage FAME_Main_Code of
 * fame_main.cpp, as they would bind the reference's (include/FAME_Main_Code.cuh:4).
 */
#include "fame_main.h"

int FAME_Main_Code(PAR Par, PROFILE* Profile) { return fame_main_code_run(&Par, Profile); }
