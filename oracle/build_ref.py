#!/usr/bin/env python3
"""Build the compiled-reference checkers into oracle/_ref/ (test infrastructure).

NcorpiON is configured at compile time through src/parameters.h.  For every
configuration below this script writes a parameters.h whose #define values are
substituted from the reference's own file into a scratch directory outside the
repository, then compiles the reference's .c files where they lie
(/root/reference/src) with the reference makefile's flags (makefile:10) and
`-I<scratch> -I- -I<reference src>`, so the generated parameters.h is found
before the original.  Nothing from the reference is copied into the
repository; only the linked executables land in oracle/_ref/<config>/:

  ref_probe               function-level probe (oracle/ref_probe.c)
  ncorpion_trace          the reference's main + call tracer (oracle/ref_trace.c)
  ncorpion_dropin_trace   same, with the GPU drop-in (needs ncorpion_amd/libncg.so)

(The drop-in NcorpiON a user builds, without the tracer, is made by
ncorpion_amd/glue/build_dropin.py.)

Run it from any directory:  python3 oracle/build_ref.py [config ...]
"""
import math
import os
import re
import shutil
import subprocess
import sys

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_ref")
SCRATCH = os.environ.get("NCG_REF_SCRATCH", "/tmp/ncg_refbuild")
CC = os.environ.get("CC", "gcc")
CFLAGS = ["-O2", "-Wall", "-Wextra", "-pedantic", "-w"]  # makefile:10 (warnings silenced)
OBJS = ["structure", "physics", "collision", "ffm", "rk4", "display", "spring"]

DT1 = 2.0 * math.pi / 256.0
DT2 = 2.0 * math.pi / 512.0

# Settings shared by the whole-run configurations (cartesian init.txt in the
# working directory, G = 1, central mass 1, no output files).
RUN_COMMON = {
    "pth": '"./"',
    "write_to_files_bool": "0",
    "central_mass_bool": "1",
    "reduce_to_COM_bool": "1",
    "random_initial_bool": "0",
    "initial_cartesian_bool": "1",
    "seed_bool": "1",
    "one_collision_only_bool": "0",
    "resume_simulation_bool": "0",
    "collision_bool": "1",
    "mutual_bool": "1",
    "falcON_bool": "0",
    "standard_tree_bool": "0",
    "elastic_collision_bool": "0",
    "inelastic_collision_bool": "1",
    "instant_merger_bool": "0",
    "fragmentation_bool": "0",
    "collision_parameter": "1.5",
    "G": "1.",
    "M_unit": "1.",
    "R_unit": "0.05",
    "disruption_threshold": "0.05",
    "high_dumping_threshold": "100.",
    "softening_parameter": "0.",
    "t_init": "0.",
}

PROBE_COMMON = {
    "pth": '"./"',
    "write_to_files_bool": "0",
    "central_mass_bool": "1",
    "random_initial_bool": "1",
    "collision_bool": "1",
    "mutual_bool": "1",
    "falcON_bool": "0",
    "standard_tree_bool": "0",
    "fragmentation_bool": "0",
    "instant_merger_bool": "0",
    "one_collision_only_bool": "0",
}


def _brute(extra):
    d = dict(extra)
    d.update({"brute_force_bool": "1", "mesh_bool": "0"})
    return d


def _mesh(extra):
    d = dict(extra)
    d.update({"brute_force_bool": "0", "mesh_bool": "1"})
    return d


def _run_cfg(kind, n, steps, dt, extra=None):
    d = dict(RUN_COMMON)
    d.update({"N_0": str(n), "N_max": str(n + 16), "time_step": repr(dt),
              "t_end": repr((steps - 0.5) * dt)})
    d.update(extra or {})
    return _brute(d) if kind == "brute" else _mesh(d)


CONFIGS = {
    # --- function-level probes (golden vectors) -------------------------------
    "probe_bf_g1": ("probe", _brute(dict(PROBE_COMMON, **{
        "G": "1.", "softening_parameter": "0.", "inelastic_collision_bool": "1",
        "elastic_collision_bool": "0", "collision_parameter": "1.5", "N_max": "40000"}))),
    "probe_bf_soft": ("probe", _brute(dict(PROBE_COMMON, **{
        "softening_parameter": "0.5", "elastic_collision_bool": "1",
        "inelastic_collision_bool": "0", "one_collision_only_bool": "1", "N_max": "5000"}))),
    "probe_bf_nocm": ("probe", _brute(dict(PROBE_COMMON, **{
        "central_mass_bool": "0", "G": "1.", "softening_parameter": "0.1",
        "inelastic_collision_bool": "1", "elastic_collision_bool": "0",
        "collision_parameter": "1.25", "N_max": "5000"}))),
    "probe_c3": ("probe", _brute(dict(PROBE_COMMON, **{
        "G": "1.", "softening_parameter": "1e-6", "collision_bool": "0",
        "inelastic_collision_bool": "0", "elastic_collision_bool": "0", "N_max": "300000"}))),
    "probe_mesh": ("probe", _mesh(dict(PROBE_COMMON, **{
        "G": "1.", "softening_parameter": "0.", "inelastic_collision_bool": "1",
        "elastic_collision_bool": "0", "collision_parameter": "1.5",
        "collision_cube_cells": "16", "collision_cube_min": "0.16",
        "how_many_neighbours": "40.0", "N_max": "5000"}))),
    "probe_mesh_el": ("probe", _mesh(dict(PROBE_COMMON, **{
        "softening_parameter": "0.25", "elastic_collision_bool": "1",
        "inelastic_collision_bool": "0", "one_collision_only_bool": "1",
        "collision_cube_cells": "32", "collision_cube_min": "0.32",
        "how_many_neighbours": "40.0", "N_max": "5000"}))),
    # Mesh module at the 2^20-body disk size (cells of side 0.01, 256 per
    # axis): CPU baseline of one mesh() / get_neighbours_mesh() call.
    "probe_c2": ("probe", _mesh(dict(PROBE_COMMON, **{
        "G": "1.", "softening_parameter": "0.", "inelastic_collision_bool": "1",
        "elastic_collision_bool": "0", "collision_parameter": "1.5",
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "N_max": "1048600"}))),
    # display() of the 2^20-body disk (CPU baseline of the output path):
    # cartesian and elliptic elements, radii and masses.
    "probe_out": ("probe", _brute(dict(PROBE_COMMON, **{
        "G": "1.", "write_to_files_bool": "1", "write_cartesian_bool": "1",
        "write_elliptic_bool": "1", "write_collisions_bool": "0", "N_max": "1048600"}))),
    # --- whole runs (drop-in parity) ------------------------------------------
    # Brute force, 2000 bodies, 200 steps of 2 pi/256, inelastic bounces.
    "run_c1": ("run", _run_cfg("brute", 2000, 200, DT1)),
    # Same, denser (larger radii in init.txt) and shorter, for many collisions.
    "run_c1_dense": ("run", _run_cfg("brute", 2000, 40, DT1)),
    # Brute force with the reference's default collision model (mergers,
    # partial and full fragmentation: bodies removed and created mid-pass),
    # and with instant mergers; 2000 bodies, 40 steps.
    "run_c1_frag": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "inelastic_collision_bool": "0", "fragmentation_bool": "1", "N_max": "4000"})),
    "run_c1_merge": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "inelastic_collision_bool": "0", "instant_merger_bool": "1"})),
    # vector_field's terms beyond the central point mass: the J2 bulge of the
    # central body (R_unit 0.05), a point-mass perturber (mass 0.01 on a
    # circular orbit of radius 5), and an inner fluid disk whose mass enters
    # mu (physics.c:79, 110-172).
    "run_c1_j2": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "J2_bool": "1", "J2_value": "0.001"})),
    "run_c1_pert": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "pert_mass": "0.01", "pert_sma": "5."})),
    "run_c1_disk": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "inner_fluid_disk_bool": "1", "inner_mass": "0.001"})),
    # J2 with the fragmentation model on the mesh path: the survivors and
    # fragments of full fragmentations (explicit term lists) start from the
    # glue's values too.
    "run_mesh_j2frag": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01",
        "J2_bool": "1", "J2_value": "0.001", "inelastic_collision_bool": "0",
        "fragmentation_bool": "1", "N_max": "40000"})),
    "run_mesh_j2": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01",
        "J2_bool": "1", "J2_value": "0.001", "pert_mass": "0.01", "pert_sma": "5."})),
    # The fragmentation model with end_of_timestep's removals: bodies closer
    # to the central body than 0.93 merge into it, bodies beyond 1.07 are
    # dropped (the centre of mass is then reduced), and the array is
    # compacted once a tenth of its slots are empty (tidy_up).
    # The fragmentation model with one collision per body and step
    # (did_collide set by every outcome, cleared by end_of_timestep).
    "run_c1_frag1": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "inelastic_collision_bool": "0", "fragmentation_bool": "1", "N_max": "4000",
        "one_collision_only_bool": "1"})),
    "run_c1_rm": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "inelastic_collision_bool": "0", "fragmentation_bool": "1", "N_max": "4000",
        "disruption_threshold": "0.93", "high_dumping_threshold": "1.07"})),
    # Mesh, 16384 bodies, cells of side 0.01 (256 per axis), 10 steps of 2 pi/512.
    # The initial mesh size comes from sma_min/sma_max/inclination_max
    # (structure.c:517), so those are set to the disk's range.
    "run_mesh": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01"})),
    # The same mesh run with the reference's default collision model
    # (mergers, partial and full fragmentation: bodies removed mid-pass,
    # fragments appended with their pairs) and with instant mergers.
    "run_mesh_frag": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01",
        "inelastic_collision_bool": "0", "fragmentation_bool": "1", "N_max": "40000"})),
    # The fragmentation model with bodies larger than the mesh size (their
    # visits scan every body, collision.c:381-414; the initial conditions
    # put them in the disk): a lower neighbour target keeps the mesh size
    # below their radii.
    "run_mesh_frag_big": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "30.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01",
        "inelastic_collision_bool": "0", "fragmentation_bool": "1", "N_max": "40000"})),
    # The fragmentation model with end_of_timestep's removals on the mesh
    # path (bodies inside 0.93 or beyond 1.07 removed, then a compaction).
    "run_mesh_rm": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01",
        "inelastic_collision_bool": "0", "fragmentation_bool": "1", "N_max": "40000",
        "disruption_threshold": "0.93", "high_dumping_threshold": "1.07"})),
    # Inelastic bounces with bodies larger than the mesh size (the
    # GPU-resident mesh loop replays such passes on the host).
    "run_mesh_big": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "30.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01"})),
    "run_mesh_merge": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01",
        "inelastic_collision_bool": "0", "instant_merger_bool": "1"})),
    # Output files (display(), every 4th step: cartesian and elliptic
    # elements, radii, masses, statistics, collision data) of a brute-force
    # run with the fragmentation model, and of a mesh run.
    "run_c1_out": ("run", _run_cfg("brute", 2000, 40, DT1, {
        "inelastic_collision_bool": "0", "fragmentation_bool": "1", "N_max": "4000",
        "write_to_files_bool": "1", "write_cartesian_bool": "1", "write_elliptic_bool": "1",
        "write_collisions_bool": "1", "output_step": "4"})),
    # Elliptic elements only (no collision file), and cartesian only with the
    # collision file: the other layouts of the files array.
    "run_c1_out_ell": ("run", _run_cfg("brute", 2000, 20, DT1, {
        "write_to_files_bool": "1", "write_cartesian_bool": "0", "write_elliptic_bool": "1",
        "write_collisions_bool": "0", "output_step": "4"})),
    "run_c1_out_cart": ("run", _run_cfg("brute", 2000, 20, DT1, {
        "inelastic_collision_bool": "0", "instant_merger_bool": "1",
        "write_to_files_bool": "1", "write_cartesian_bool": "1", "write_elliptic_bool": "0",
        "write_collisions_bool": "1", "output_step": "4"})),
    "run_mesh_out": ("run", _run_cfg("mesh", 16384, 10, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01",
        "write_to_files_bool": "1", "write_cartesian_bool": "1", "write_elliptic_bool": "1",
        "output_step": "2"})),
    # Full-size mesh disk of 2^20 bodies (CPU baseline of the mesh module).
    "run_c2": ("run", _run_cfg("mesh", 1048576, 3, DT2, {
        "collision_cube_cells": "256", "collision_cube_min": "2.56",
        "how_many_neighbours": "100.0", "switch_to_brute_force": "110",
        "sma_min": "0.9", "sma_max": "1.1", "inclination_max": "0.01"})),
}


def write_parameters(dst, overrides):
    src = open(os.path.join(REF_SRC, "parameters.h")).read()
    seen = set()

    def sub(m):
        name = m.group(1)
        if name in overrides:
            seen.add(name)
            return "#define %s %s" % (name, overrides[name])
        return m.group(0)

    text = re.sub(r"^#define\s+(\w+)\s+[^\n]*$", sub, src, flags=re.M)
    missing = set(overrides) - seen
    if missing:
        raise SystemExit("parameters.h has no #define for: %s" % ", ".join(sorted(missing)))
    os.makedirs(dst, exist_ok=True)
    with open(os.path.join(dst, "parameters.h"), "w") as f:
        f.write(text)


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise SystemExit("build failed")


def build(name):
    kind, overrides = CONFIGS[name]
    scratch = os.path.join(SCRATCH, name)
    if os.path.isdir(scratch):
        shutil.rmtree(scratch)
    write_parameters(scratch, overrides)
    inc = ["-I" + scratch, "-I-", "-I" + REF_SRC]
    objs = []
    for o in OBJS:
        obj = os.path.join(scratch, o + ".o")
        run([CC] + CFLAGS + inc + ["-c", os.path.join(REF_SRC, o + ".c"), "-o", obj])
        objs.append(obj)
    out = os.path.join(OUT, name)
    os.makedirs(out, exist_ok=True)
    if kind == "probe":
        run([CC] + CFLAGS + inc + [os.path.join(HERE, "ref_probe.c")] + objs +
            ["-Wl,--wrap=collision_treatment", "-lm", "-o", os.path.join(out, "ref_probe")])
        return
    main_o = os.path.join(scratch, "main.o")
    run([CC] + CFLAGS + inc + ["-c", os.path.join(REF_SRC, "main.c"), "-o", main_o])
    wraps = ("-Wl,--wrap=vector_field,--wrap=brute_force,--wrap=mesh,--wrap=get_neighbours_mesh,"
             "--wrap=display")
    trace_wraps = wraps + ",--wrap=collision_treatment,--wrap=three_largest_moonlets"
    run([CC] + CFLAGS + inc + [os.path.join(HERE, "ref_trace.c"), main_o] + objs +
        [trace_wraps, "-lm", "-o", os.path.join(out, "ncorpion_trace")])
    # The glue also records the slots fragmentation creates (get_free_index).
    wraps += ",--wrap=get_free_index"
    trace_wraps += ",--wrap=get_free_index"
    lib_dir = os.path.join(REPO, "ncorpion_amd")
    if not os.path.exists(os.path.join(lib_dir, "libncg.so")):
        print("  (libncg.so not built: skipping drop-in executables)")
        return
    glue = os.path.join(REPO, "ncorpion_amd", "glue")
    ginc = inc + ["-I" + os.path.join(REPO, "include")]
    glue_o = os.path.join(scratch, "ncg_dropin.o")
    run([CC] + CFLAGS + ginc + ["-c", os.path.join(glue, "ncg_dropin.c"), "-o", glue_o])
    # The executables find libncg.so relative to themselves.
    rpath = "-Wl,-rpath,$ORIGIN/../../../ncorpion_amd"
    link = ["-L" + lib_dir, "-lncg", rpath, "-lm"]
    run([CC] + CFLAGS + inc + ["-DTRACE_DROPIN", os.path.join(HERE, "ref_trace.c"), glue_o, main_o]
        + objs + [trace_wraps] + link + ["-o", os.path.join(out, "ncorpion_dropin_trace")])


def main(argv):
    if not os.path.isdir(REF_SRC):
        print("build_ref: %s not present; using prebuilt oracle/_ref" % REF_SRC)
        return 0
    names = argv or list(CONFIGS)
    for n in names:
        print("build_ref:", n)
        build(n)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
