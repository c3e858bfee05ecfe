#!/usr/bin/env python3
"""Build the reference openQCD-2.4.2 hot-path modules into driver programs.

Test infrastructure only.  Compiles the reference's own C sources where they
lie under /root/reference (nothing is copied into this repository) together
with oracle/ref_driver.c, one executable per lattice geometry, into
oracle/_ref/ref_<tag>.  The lattice sizes and thread-block sizes are
compile-time constants of the reference (include/global.h:18-36); for each
geometry a modified global.h is generated into oracle/_ref/obj/<tag>/ and
force-included, so that its include guard shadows the original.

The reference's OpenMP-offload annotations (include/su3v.h) need an OpenMP 5
front end, so the ROCm clang is used; MPI comes from /opt/conda (NPROC=1,
singleton MPI_Init).  Usage:  python3 oracle/build_ref.py [tag ...]
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
OUT = os.path.join(HERE, "_ref")
CC = "/opt/rocm/llvm/bin/clang"
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib"

MODULES = """block/block block/blk_grid block/map_u2blk block/map_sw2blk block/map_s2blk
dirac/Dw_dble dirac/Dw dirac/Dw_bnd dfl/dfl_geometry flags/action_parms flags/dfl_parms
flags/flags flags/hmc_parms flags/lat_parms flags/rat_parms flags/sap_parms flags/smd_parms
lattice/bcnds lattice/blk_geometry lattice/ftidx lattice/geogen lattice/geometry lattice/uidx
linalg/salg linalg/salg_dble linalg/liealg linalg/cmatrix_dble linsolv/fgcr random/ranlux
random/ranlux_common random/ranlxs random/ranlxd random/gauss sap/blk_solv sap/sap_com
sap/sap sap/sap_gcr sflds/Pbnd sflds/Pbnd_dble sflds/sflds sflds/scom sflds/sdcom
sflds/unorm su3fcts/chexp su3fcts/su3prod su3fcts/su3ren su3fcts/cm3x3 su3fcts/random_su3
sw_term/pauli sw_term/pauli_dble sw_term/swalg sw_term/swexp sw_term/swflds sw_term/sw_term
tcharge/ftcom tcharge/ftensor uflds/plaq_sum uflds/shift uflds/udcom uflds/uflds
uflds/uinit utils/endian utils/error utils/mutils utils/qsum utils/utils utils/wspace
forces/force0 mdflds/mdflds mdflds/fcom wflow/wflow uflds/bstap archive/archive archive/sarchive
dfl/dfl_subspace dfl/dfl_modes dfl/dfl_sap_gcr dfl/ltl_gcr little/Aw little/Aw_blk little/Aw_com1
little/Aw_com2 little/Aw_dble little/Aw_gen little/Aw_hop little/Aw_ops little/ltl_modes vflds/vcom
vflds/vdcom vflds/vflds vflds/vinit linalg/cmatrix linalg/valg linalg/valg_dble linsolv/fgcr4vd
linsolv/gcr4v""".split()

# tag: (L[4], Ltrd[4], simd)   simd=False -> portable C branches (parity)
#                              simd=True  -> -DAVX -DFMA3 -DPM (timing)
CONFIGS = {
    "8x8x8x8_t8": ((8, 8, 8, 8), (8, 8, 8, 8), False),
    "8x8x8x8_t4": ((8, 8, 8, 8), (4, 4, 4, 4), False),
    "16x8x8x8_t8": ((16, 8, 8, 8), (8, 8, 8, 8), False),
    # generic SAP block sizes: 8^4 blocks on 16^4 and 6^4 blocks on 12^4
    # (16 blocks, an even multiple of NTHREAD = 8 and 4)
    "16x16x16x16_t8": ((16, 16, 16, 16), (8, 8, 8, 16), False),
    "12x12x12x12_t4": ((12, 12, 12, 12), (6, 6, 12, 12), False),
    "64x32x32x32_avx": ((64, 32, 32, 32), (16, 16, 16, 32), True),
    # SAP+GCR CPU baselines: 8^4 (4^4 blocks need NTHREAD <= 8) and 48^4
    "8x8x8x8_avx": ((8, 8, 8, 8), (8, 4, 4, 4), True),
    "48x48x48x48_avx": ((48, 48, 48, 48), (24, 24, 24, 24), True),
}


def global_h(tag):
    L, T, _ = CONFIGS[tag]
    src = open(os.path.join(REF, "include", "global.h")).read()
    rep = {"NPROC0": 1, "NPROC1": 1, "NPROC2": 1, "NPROC3": 1}
    for mu in range(4):
        rep["L%d" % mu] = L[mu]
        rep["L%d_TRD" % mu] = T[mu]
    for k, v in rep.items():
        src, n = re.subn(r"(?m)^#define %s\s+\S+\s*$" % k, "#define %s %d" % (k, v), src)
        assert n == 1, k
    d = os.path.join(OUT, "obj", tag)
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "global.h")
    if not os.path.exists(p) or open(p).read() != src:
        open(p, "w").write(src)
    return p


def build(tag, jobs=8):
    if not os.path.isdir(REF):
        print("reference tree not present; skipping", tag)
        return None
    gh = global_h(tag)
    simd = CONFIGS[tag][2]
    objdir = os.path.dirname(gh)
    # the reference headers first: the MPI include directory also holds an
    # unrelated archive.h
    flags = ["-std=gnu11", "-O2", "-fopenmp", "-w", "-include", gh,
             "-I" + os.path.join(REF, "include"), "-I" + MPI_INC]
    if simd:
        flags += ["-DAVX", "-DFMA3", "-DPM", "-mavx2", "-mfma"]
    else:
        # portable C branches, evaluated without contraction as gcc -O does
        flags += ["-ffp-contract=off"]

    stamp = os.path.join(objdir, "flags.txt")
    key = " ".join(flags)
    if not os.path.exists(stamp) or open(stamp).read() != key:
        for f in os.listdir(objdir):
            if f.endswith(".o"):
                os.remove(os.path.join(objdir, f))
        open(stamp, "w").write(key)

    def cc(src, obj):
        if os.path.exists(obj) and os.path.getmtime(obj) > max(
                os.path.getmtime(src), os.path.getmtime(gh)):
            return None
        # the driver owns the global arrays (global.h, MAIN_PROGRAM branch)
        pre = ["-DMAIN_PROGRAM", "-include", "stddef.h"] if src.endswith("ref_driver.c") else []
        r = subprocess.run([CC] + pre + flags + ["-c", src, "-o", obj],
                           capture_output=True, text=True)
        return None if r.returncode == 0 else (src, r.stderr[-2000:])

    work = [(os.path.join(REF, "modules", m + ".c"),
             os.path.join(objdir, os.path.basename(m) + ".o")) for m in MODULES]
    work.append((os.path.join(HERE, "ref_driver.c"), os.path.join(objdir, "ref_driver.o")))
    with ThreadPoolExecutor(jobs) as ex:
        errs = [e for e in ex.map(lambda w: cc(*w), work) if e]
    if errs:
        for src, msg in errs:
            print("FAILED", src, msg)
        raise SystemExit(1)
    exe = os.path.join(OUT, "ref_" + tag)
    objs = [w[1] for w in work]
    # With -DAVX the fork's su3fcts/cm3x3.c leaves out cm3x3_unityv, which
    # uflds/uinit.c:set_ud2unityv references; that function is never called
    # on the timed path, so the reference is linked with the reference left
    # unresolved (no replacement is supplied).
    extra = ["-Wl,--unresolved-symbols=ignore-in-object-files"] if simd else []
    r = subprocess.run([CC, "-fopenmp", "-o", exe] + objs + extra +
                       ["-L" + MPI_LIB, "-Wl,-rpath," + MPI_LIB, "-lmpi", "-lm"],
                       capture_output=True, text=True)
    if r.returncode:
        print(r.stderr[-3000:])
        raise SystemExit(1)
    return exe


def main(argv):
    tags = argv or list(CONFIGS)
    for t in tags:
        exe = build(t)
        if exe:
            print("built", exe)


if __name__ == "__main__":
    main(sys.argv[1:])
