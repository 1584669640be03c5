"""Builds the reference's own MDSTEDC chain (its Fortran sources, where they lie under
/root/reference/SRC) with oracle/ref_drive.f90 into oracle/_ref/libhssref.so.  Test
infrastructure only: the tests use it to check the C oracle against the reference.

The reference calls BLAS and LAPACK (DGEMM, DLAED4, DSTEQR, ...).  The only LAPACK in
this image is the OpenBLAS that scipy ships (scipy.libs/libscipy_openblas-*.so), which
exports every routine under a "scipy_" prefix.  The recipe compiles the sources with
flang, renames each undefined BLAS/LAPACK symbol of the objects to its scipy_ name
(objcopy --redefine-syms), and links against that library.  Two more renames hook the
per-merge trace (ref_drive.f90): mdlaed1's MDLAED2 call and mdlaed2's DROT.

Run: python3 oracle/build_ref.py   (does nothing when /root/reference is absent)
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/SRC"
OUT = os.path.join(HERE, "_ref")
FC = "/opt/rocm/llvm/bin/flang"
# module order matters: a file's USEd modules come first
SOURCES = ["comp_Wtime", "aux_hss", "BasicMM", "ConstructHssd", "CauchyHssEig", "CauchyHssvd_VPS",
           "CauchyHss_VP2", "mdlaed0", "mdlaed1", "mdlaed2", "mdlaed3", "mdstedc",
           "mdbdsdc", "mdlasd0", "mdlasd1", "mdlasd2", "mdlasd3", "mdlasd8",
           "mdsbevd", "mdsbed1", "mdsbed2", "mdsbed22", "mdsbed3", "mdsbed4", "mdsbed5",
           "DHSSMM", "cauchyhssmm", "Cauchylike", "SFtransf"]
HOOKS = {"mdlaed1": {"mdlaed2_": "refw_mdlaed2_"}, "mdlaed2": {"drot_": "refw_drot_"},
         "mdlasd1": {"mdlasd2_": "refw_mdlasd2_"}, "mdlasd2": {"drot_": "refw_drot_all_"},
         "mdsbed1": {"mdsbed2_": "refw_mdsbed2_", "mdsbed22_": "refw_mdsbed22_"},
         "mdsbed2": {"drot_": "refw_drot_q_"}, "mdsbed22": {"drot_": "refw_drot_q_"}}
# The row pivot search's early-exit test (CauchyPivtEig, CauchyHssEig.f90:676 and :678)
# forms its pivot with F(PL(k)): F is the node's local row-generator array (length M) but
# PL(k) is a global pole index, so the test reads whatever that local buffer holds at
# that index (stale values of earlier nodes in Cauchy2hssEig_omp, which allocates it with
# K entries).  The column search (:829, :831) reads FF(PL(k)).  A second library,
# libhssref_fix.so, is built from the same sources with those two expressions reading FF;
# the tests show the oracle agrees with it exactly and with the unmodified reference up
# to the pivots that stale read changes.
FIX = {"CauchyHssEig": [("( F(PL(k))-FF(PU(k)) )", "( FF(PL(k))-FF(PU(k)) )"),
                        ("( F(PL(k))-FF(PU(k)+1) )", "( FF(PL(k))-FF(PU(k)+1) )")]}


def openblas():
    import scipy
    libs = glob.glob(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs",
                                  "libscipy_openblas*.so"))
    if not libs:
        raise RuntimeError("scipy's OpenBLAS not found")
    return libs[0]


def exported(lib):
    out = subprocess.check_output(["nm", "-D", "--defined-only", lib], text=True)
    return {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[-1].startswith("scipy_")}


def undefined(obj):
    out = subprocess.check_output(["nm", "-u", obj], text=True)
    return [ln.split()[-1] for ln in out.splitlines() if ln.strip()]


def run(cmd, verbose):
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)


def build(verbose=False):
    if not os.path.isdir(REF):
        return None
    lib = openblas()
    exp = exported(lib)
    obj_dir = os.path.join(OUT, "obj")
    mod_dir = os.path.join(OUT, "mod")
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(mod_dir, exist_ok=True)
    # -O0: flang's optimised MDSBEVD enters its merge loop once more with one subproblem
    # left (label 80's IF (SUBPBS.GT.1) taken at SUBPBS = 1, mdsbevd.f90:405-456) and
    # calls MDSBED1 with CUTPNT = N, which returns INFO = -7; unoptimised code runs the
    # loop as written.  No fast-math and no FMA contraction at any level, so the
    # arithmetic is the same.
    flags = ["-O0", "-fPIC", "-module-dir", mod_dir, "-I", mod_dir]
    objs = []
    for name in SOURCES + ["ref_drive"]:
        src = os.path.join(HERE, name + ".f90") if name == "ref_drive" else os.path.join(REF, name + ".f90")
        if not os.path.exists(src):
            src = os.path.join(REF, name + ".f")  # fixed-form sources
        obj = os.path.join(obj_dir, name + ".o")
        run([FC] + flags + ["-c", src, "-o", obj], verbose)
        ren = {}
        for s in undefined(obj):
            if "scipy_" + s in exp:
                ren[s] = "scipy_" + s
        ren.update(HOOKS.get(name, {}))
        if ren:
            mp = os.path.join(obj_dir, name + ".syms")
            with open(mp, "w") as fh:
                for a, b in sorted(ren.items()):
                    fh.write("%s %s\n" % (a, b))
            run(["objcopy", "--redefine-syms=" + mp, obj], verbose)
        objs.append(obj)
    libdir = os.path.dirname(lib)

    def link(so, objs):
        run([FC, "-shared", "-o", so + ".tmp"] + objs +
            ["-L" + libdir, "-l:" + os.path.basename(lib), "-Wl,-rpath," + libdir, "-Wl,--no-undefined"], verbose)
        os.replace(so + ".tmp", so)

    so = os.path.join(OUT, "libhssref.so")
    link(so, objs)
    # the corrected variant: the fixed source is written under oracle/_ref only
    fixed = list(objs)
    for name, subs in FIX.items():
        with open(os.path.join(REF, name + ".f90")) as fh:
            text = fh.read()
        for a, b in subs:
            if text.count(a) != 1:
                raise RuntimeError("expected one %r in %s.f90" % (a, name))
            text = text.replace(a, b)
        src = os.path.join(OUT, name + "_fix.f90")
        with open(src, "w") as fh:
            fh.write(text)
        obj = os.path.join(obj_dir, name + "_fix.o")
        fmod = os.path.join(OUT, "mod_fix")
        os.makedirs(fmod, exist_ok=True)
        run([FC, "-O0", "-fPIC", "-module-dir", fmod, "-I", mod_dir, "-c", src, "-o", obj], verbose)
        run(["objcopy", "--redefine-syms=" + os.path.join(obj_dir, name + ".syms"), obj], verbose)
        fixed[SOURCES.index(name)] = obj
    link(os.path.join(OUT, "libhssref_fix.so"), fixed)
    return so


if __name__ == "__main__":
    p = build(verbose=True)
    print(p or "no reference sources: nothing built")
    sys.exit(0)
