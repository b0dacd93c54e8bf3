"""Regenerates the golden fixtures in this folder from the reference library.

The reference library is compiled from the reference's own C sources by
``oracle/Makefile`` (target ``ref``); this script compiles ``gen_golden.c``
against the reference's header, links it to the serial and the OpenMP builds,
and runs both.  It needs the reference source tree and is run by hand, in the
container that has it; the fixtures it writes are committed.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFSRC = os.environ.get("REFSRC", "/root/reference/src")


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"])
    refdir = os.path.join(ROOT, "oracle", "_ref")
    for flavour in ("serial", "omp"):
        exe = os.path.join(refdir, "gen_golden_" + flavour)
        cmd = ["gcc", "-O2", "-std=c99", "-D_GNU_SOURCE", "-I" + REFSRC,
               "-o", exe, os.path.join(HERE, "gen_golden.c"),
               "-L" + os.path.join(refdir, flavour), "-lrebound",
               "-Wl,-rpath," + os.path.join(refdir, flavour), "-lm"]
        if flavour == "omp":
            cmd.append("-fopenmp")
        subprocess.check_call(cmd)
        env = dict(os.environ, OMP_NUM_THREADS="4")
        subprocess.check_call([exe, HERE, flavour], env=env)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
