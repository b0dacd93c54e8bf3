"""Regenerates the golden fixtures in this folder from the reference library.

The reference library is compiled from the reference's own C sources by
``oracle/Makefile`` (target ``ref``); this script compiles ``gen_golden.c``
against the reference's header, links it to the serial and the OpenMP builds,
and runs both.  It needs the reference source tree and is run by hand, in the
container that has it; the fixtures it writes are committed.

Full-size states are too large to commit: the serial run writes them as raw
records (x y z vx vy vz of body 0..N-1, in order) and this script stores the
SHA-256 of each record instead.
"""
import hashlib
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFSRC = os.environ.get("REFSRC", "/root/reference/src")


# raw file -> (fixture, bodies per record)
RAW = {"plummer65536_seed42.raw": ("plummer65536_seed42_sha256.npy", 65536),
       "sheet200k.raw": ("sheet200k_sha256.npy", 200000)}


def hash_records(rawdir):
    for raw, (fixture, n) in RAW.items():
        data = np.fromfile(os.path.join(rawdir, raw), dtype="<f8").reshape(-1, n, 6)
        digests = np.array([np.frombuffer(hashlib.sha256(np.ascontiguousarray(rec).tobytes()).digest(),
                                          dtype=np.uint8) for rec in data])
        np.save(os.path.join(HERE, fixture), digests)
        os.remove(os.path.join(rawdir, raw))


def main():
    # `python make_golden.py megno` writes only the variational/MEGNO fixtures
    only = sys.argv[1:2]
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
        rawdir = os.path.join(refdir, "raw")
        if flavour == "serial":
            os.makedirs(rawdir, exist_ok=True)
            for f in os.listdir(rawdir):
                os.remove(os.path.join(rawdir, f))
            env["GEN_GOLDEN_RAWDIR"] = rawdir
        subprocess.check_call([exe, HERE, flavour] + only, env=env)
        if flavour == "serial" and not only:
            hash_records(rawdir)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
