"""Writes tests/golden/pi.json: the digit strings of preclib's pi program.

Run in a container that has GMP and gcc; the GPU box does not need either.
gen_pi_gmp.c evaluates pi_chudnovsky.cpp's exact integer formula with GMP and
writes the same file the reference writes with --file (the digit string and a
newline, pi_chudnovsky.cpp:223-230).  Where the reference's own program has
been built from its sources (oracle/Makefile, target ref_pi) and its output
for a digit count exists in _work/pi_ref_<d>.txt, the two files are compared
byte for byte and the result recorded.

Entry: digits -> sha256 of the file bytes, length, first and last 100 bytes
of the digit string.
"""
import hashlib
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
WORK = os.path.join(ROOT, "_work")
DIGITS = [0, 1, 2, 9, 10, 100, 1000, 12345, 100000, 1000000, 10000000]


def main():
    os.makedirs(WORK, exist_ok=True)
    exe = os.path.join(WORK, "gen_pi_gmp")
    subprocess.check_call(["gcc", "-O2", "-I/opt/conda/include", "-o", exe, os.path.join(HERE, "gen_pi_gmp.c"),
                           "-L/opt/conda/lib", "-lgmp", "-Wl,-rpath,/opt/conda/lib"])
    out = {"format": "sha256 of the --file output (digit string + newline)", "pi": []}
    for d in DIGITS:
        path = os.path.join(WORK, "pi_gmp_%d.txt" % d)
        subprocess.check_call([exe, str(d), path])
        data = open(path, "rb").read()
        s = data.rstrip(b"\n").decode()
        ref = os.path.join(WORK, "pi_ref_%d.txt" % d)
        match = None
        if os.path.exists(ref):
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
