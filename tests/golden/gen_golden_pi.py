"""Writes tests/golden/pi.json: the digit strings of preclib's pi program.

Run in a container that has GMP, gcc and the reference tree; the GPU box
needs none of them.  Two independent sources, compared byte for byte:

* gen_pi_gmp.c evaluates pi_chudnovsky.cpp's exact integer formula with GMP
  and writes the file the reference writes with --file (the digit string
  and a newline, pi_chudnovsky.cpp:223-230);
* the reference's own pi program, built from its sources by oracle/Makefile
  (target ref_pi, into oracle/_ref/pi_chudnovsky), run with DIGITS --file.

reference_program_match records the comparison (null where the reference
run was skipped: --ref-max caps the digit counts it is run for, since it
takes about two minutes at 10^7 digits on one core).  Everything is built
and written in a temporary directory.

Entry: digits -> sha256 of the file bytes, length, first and last 100 bytes
of the digit string.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DIGITS = [0, 1, 2, 9, 10, 100, 1000, 12345, 100000, 1000000, 10000000]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref-max", type=int, default=10 ** 7, help="largest digit count to run the reference for")
    args = ap.parse_args()
    ref_exe = os.path.join(ROOT, "oracle", "_ref", "pi_chudnovsky")
    subprocess.check_call(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile"), "ref_pi"])
    have_ref = os.path.exists(ref_exe)
    out = {"format": "sha256 of the --file output (digit string + newline)", "pi": []}
    with tempfile.TemporaryDirectory(prefix="gen_pi_") as tmp:
        exe = os.path.join(tmp, "gen_pi_gmp")
        subprocess.check_call(["gcc", "-O2", "-I/opt/conda/include", "-o", exe, os.path.join(HERE, "gen_pi_gmp.c"),
                               "-L/opt/conda/lib", "-lgmp", "-Wl,-rpath,/opt/conda/lib"])
        for d in DIGITS:
            path = os.path.join(tmp, "pi_gmp_%d.txt" % d)
            subprocess.check_call([exe, str(d), path])
            data = open(path, "rb").read()
            s = data.rstrip(b"\n").decode()
            match = None
            if have_ref and d <= args.ref_max:
                ref = os.path.join(tmp, "pi_ref_%d.txt" % d)
                subprocess.run([ref_exe, str(d), "--file", ref], check=True, capture_output=True)
                match = open(ref, "rb").read() == data
                if not match:
                    sys.exit("reference program output differs at d=%d" % d)
            out["pi"].append({"digits": d, "sha256": hashlib.sha256(data).hexdigest(), "len": len(s),
                              "head": s[:100], "tail": s[-100:], "reference_program_match": match})
            print(d, len(s), match, flush=True)
    with open(os.path.join(HERE, "pi.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote pi.json")


if __name__ == "__main__":
    main()
