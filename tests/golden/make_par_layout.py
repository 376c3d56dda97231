"""Writes tests/golden/par_layout.json: sizeof / offsetof of the call-surface
structs (PAR and its members, PROFILE, LS, ES, VERTEX) as the reference's own
header lays them out (include/FAME_Internal_Common.h under the reference tree),
and the linker symbol of the reference's C++ declaration of FAME_Main_Code
(include/FAME_Main_Code.cuh:4), read with nm from an object compiled here.

A small program of our own that includes the reference header is compiled with
g++ into a temporary directory and run; nothing of the reference is copied.
tests/test_layout.py compiles the same program against include/fame_main.h and
compares.  Run from the repository root (needs the reference tree):

    python tests/golden/make_par_layout.py
"""
import json
import os
import subprocess
import sys
import tempfile

REF = os.environ.get("FAME_REFERENCE_DIR", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

# (struct, members); every member's offset and the struct's size are tabulated
FIELDS = [
    ("PAR", ["flag", "mesh", "lattice", "material", "recip_lattice", "ls", "es"]),
    ("FLAG", ["device", "printf_user_option", "printf_parameter", "create_parameter", "create_B_inout",
              "create_wave_vector", "save_eigen_vector", "grid_nums_max", "mem_size", "radius_adjustment",
              "radius"]),
    ("MESH", ["grid_nums", "edge_len", "mesh_len"]),
    ("LATTICE", ["lattice_type", "Permutation", "lattice_vec_a", "lattice_vec_a_orig", "Omega",
                 "lattice_constant"]),
    ("LATTICE_CONSTANT", ["a", "b", "c", "alpha", "beta", "gamma", "theta_1", "theta_2", "theta_3", "m1", "m2",
                          "m3", "m4", "t1", "t2", "t3", "t4", "flag", "length_a1", "length_a2", "length_a3"]),
    ("MATERIAL", ["data_name", "material_type", "Binout", "material_num", "sphere_num", "sphere_centers",
                  "sphere_radius", "cylinder_num", "cylinder_top_centers", "cylinder_bot_centers",
                  "cylinder_radius", "num_ele_permitt_in", "ele_permitt_in", "ele_permitt_out",
                  "flag_radius_adjustment", "flag_radius"]),
    ("RECIP_LATTICE", ["part_num", "Wave_vec_num", "WaveVector", "path_string", "reciprocal_lattice_vector_b",
                       "vertex"]),
    ("VERTEX", [c for c in "ABCDEFGHIJKLMNOPQRSTUVWXYZ"] + ["s", "t", "k"]),
    ("LS", ["maxit", "tol"]),
    ("ES", ["nwant", "nstep", "maxit", "tol"]),
    ("PROFILE", ["idx", "ls_iter", "es_iter", "ls_time", "es_time"]),
]


def layout_source(include_line):
    lines = [include_line, "#include <stdio.h>", "#include <stddef.h>", "int main(void) {", '  printf("{");']
    first = True
    for st, members in FIELDS:
        sep = "" if first else ","
        first = False
        lines.append(f'  printf("{sep}\\"{st}\\": {{\\"sizeof\\": %zu", sizeof({st}));')
        for m in members:
            lines.append(f'  printf(", \\"{m}\\": %zu", offsetof({st}, {m}));')
        lines.append('  printf("}");')
    lines += ['  printf("}\\n");', "  return 0;", "}"]
    return "\n".join(lines) + "\n"


def run_layout(compiler, include_dir, include_line, lang_flags, workdir):
    src = os.path.join(workdir, "layout" + (".cpp" if "++" in compiler else ".c"))
    exe = os.path.join(workdir, "layout_" + os.path.basename(compiler).replace("+", "p"))
    with open(src, "w") as f:
        f.write(layout_source(include_line))
    subprocess.run([compiler, *lang_flags, "-w", "-I", include_dir, src, "-o", exe], check=True)
    return json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)


def main():
    inc = os.path.join(REF, "include")
    if not os.path.isdir(inc):
        sys.exit(f"reference tree not found at {REF}")
    with tempfile.TemporaryDirectory() as td:
        layout = run_layout("g++", inc, '#include "FAME_Internal_Common.h"', ["-std=gnu++11"], td)
        caller = os.path.join(td, "caller.cpp")
        with open(caller, "w") as f:
            f.write('#include "FAME_Internal_Common.h"\n#include "FAME_Main_Code.cuh"\n'
                    "int call(PAR p, PROFILE* q) { return FAME_Main_Code(p, q); }\n")
        obj = os.path.join(td, "caller.o")
        subprocess.run(["g++", "-std=gnu++11", "-w", "-I", inc, "-c", caller, "-o", obj], check=True)
        nm = subprocess.run(["nm", "-u", obj], check=True, capture_output=True, text=True).stdout
        syms = [ln.split()[-1] for ln in nm.splitlines() if "FAME_Main_Code" in ln]
    out = {"generator": "tests/golden/make_par_layout.py", "source": "include/FAME_Internal_Common.h, "
           "include/FAME_Main_Code.cuh (reference tree), g++ x86-64", "layout": layout,
           "FAME_Main_Code_symbol": syms[0] if syms else None}
    with open(os.path.join(HERE, "par_layout.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"PAR {layout['PAR']['sizeof']} bytes, symbol {out['FAME_Main_Code_symbol']}")


if __name__ == "__main__":
    main()
