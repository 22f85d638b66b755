#!/usr/bin/env python3
"""Build drop-in NcorpiON executables: NcorpiON's own sources, configured
through its parameters.h, linked with this glue and libncg.so.

This is the build a NcorpiON user makes (the make recipe of the integration
notes), applied to the benchmark workloads.  For each configuration a
parameters.h is written into a scratch directory outside the repository, with
#define values substituted into the reference's own file; the reference's .c
files are compiled where they lie with its makefile's flags (makefile:10) and
`-I<scratch> -I- -I<reference src>`, so the written parameters.h is found
first.  Nothing from NcorpiON is copied into the repository; only the linked
executables land in ncorpion_amd/glue/_build/<config>/ncorpion_gpu.

Configurations:
  c1  brute force, 2000 bodies, 200 steps of 2 pi/256, inelastic bounces f = 1.5
  c2  mesh, 2^20 bodies, cells of side 0.01 (256 per axis), 3 steps of 2 pi/512,
      inelastic bounces f = 1.5

The executables read init.txt (cartesian, G = 1, central mass 1) from their
working directory and take the glue's environment switches (NCG_PRECISION,
NCG_TIMING, NCG_DEVICE).

    python3 ncorpion_amd/glue/build_dropin.py [config ...]
"""
import math
import os
import re
import shutil
import subprocess
import sys

NCORPION_SRC = os.environ.get("NCORPION_SRC", "/root/reference/src")
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
SCRATCH = os.environ.get("NCG_DROPIN_SCRATCH", "/tmp/ncg_dropin_build")
CC = os.environ.get("CC", "gcc")
CFLAGS = ["-O2", "-Wall", "-Wextra", "-pedantic", "-w"]
UNITS = ["structure", "physics", "collision", "ffm", "rk4", "display", "spring", "main"]
WRAPS = ("-Wl,--wrap=vector_field,--wrap=brute_force,--wrap=mesh,--wrap=get_neighbours_mesh,"
         "--wrap=get_free_index,--wrap=display")


def _settings(n, steps, dt, module):
    s = {
        "pth": '"./"', "write_to_files_bool": "0", "central_mass_bool": "1",
        "reduce_to_COM_bool": "1", "random_initial_bool": "0", "initial_cartesian_bool": "1",
        "seed_bool": "1", "resume_simulation_bool": "0", "collision_bool": "1",
        "mutual_bool": "1", "falcON_bool": "0", "standard_tree_bool": "0",
        "one_collision_only_bool": "0", "elastic_collision_bool": "0",
        "inelastic_collision_bool": "1", "instant_merger_bool": "0", "fragmentation_bool": "0",
        "collision_parameter": "1.5", "G": "1.", "M_unit": "1.", "R_unit": "0.05",
        "disruption_threshold": "0.05", "high_dumping_threshold": "100.",
        "softening_parameter": "0.", "t_init": "0.",
        "N_0": str(n), "N_max": str(n + 16), "time_step": repr(dt),
        "t_end": repr((steps - 0.5) * dt),
        "brute_force_bool": "1" if module == "brute" else "0",
        "mesh_bool": "1" if module == "mesh" else "0",
    }
    if module == "mesh":
        s.update({"collision_cube_cells": "256", "collision_cube_min": "2.56",
                  "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
                  "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01"})
    return s


CONFIGS = {
    "c1": _settings(2000, 200, 2.0 * math.pi / 256.0, "brute"),
    "c2": _settings(1 << 20, 3, 2.0 * math.pi / 512.0, "mesh"),
}


def write_parameters(dst, settings):
    text = open(os.path.join(NCORPION_SRC, "parameters.h")).read()
    used = set()

    def sub(m):
        if m.group(1) in settings:
            used.add(m.group(1))
            return "#define %s %s" % (m.group(1), settings[m.group(1)])
        return m.group(0)

    text = re.sub(r"^#define\s+(\w+)\s+[^\n]*$", sub, text, flags=re.M)
    if set(settings) - used:
        raise SystemExit("parameters.h has no #define for: %s"
                         % ", ".join(sorted(set(settings) - used)))
    os.makedirs(dst, exist_ok=True)
    with open(os.path.join(dst, "parameters.h"), "w") as f:
        f.write(text)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise SystemExit("build failed")


def build(name):
    lib_dir = os.path.join(REPO, "ncorpion_amd")
    if not os.path.exists(os.path.join(lib_dir, "libncg.so")):
        raise SystemExit("libncg.so is not built (make -C ncorpion_amd/csrc)")
    scratch = os.path.join(SCRATCH, name)
    if os.path.isdir(scratch):
        shutil.rmtree(scratch)
    write_parameters(scratch, CONFIGS[name])
    inc = ["-I" + scratch, "-I-", "-I" + NCORPION_SRC]
    objs = []
    for u in UNITS:
        o = os.path.join(scratch, u + ".o")
        _run([CC] + CFLAGS + inc + ["-c", os.path.join(NCORPION_SRC, u + ".c"), "-o", o])
        objs.append(o)
    glue_o = os.path.join(scratch, "ncg_dropin.o")
    _run([CC] + CFLAGS + inc + ["-I" + os.path.join(REPO, "include"), "-c",
                                os.path.join(HERE, "ncg_dropin.c"), "-o", glue_o])
    wrap_o = os.path.join(scratch, "ncg_wrap.o")
    _run([CC] + CFLAGS + ["-c", os.path.join(HERE, "ncg_wrap.c"), "-o", wrap_o])
    out = os.path.join(OUT, name)
    os.makedirs(out, exist_ok=True)
    # The executable finds libncg.so relative to itself.
    _run([CC] + CFLAGS + objs + [glue_o, wrap_o, WRAPS, "-L" + lib_dir, "-lncg",
                                 "-Wl,-rpath,$ORIGIN/../../..", "-lm", "-o",
                                 os.path.join(out, "ncorpion_gpu")])


def main(argv):
    if not os.path.isdir(NCORPION_SRC):
        print("build_dropin: %s not present; using the prebuilt executables" % NCORPION_SRC)
        return 0
    for name in argv or list(CONFIGS):
        print("build_dropin:", name)
        build(name)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
