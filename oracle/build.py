"""Compiles the C restatement of the oracle (oracle/lbm_oracle_c.c) into
oracle/_build/liblbm_oracle_c.so -- test infrastructure only.

gcc -O2 for x86-64 without -march: no FMA contraction, IEEE double arithmetic.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liblbm_oracle_c.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas",
                    "-o", tmp, os.path.join(HERE, "lbm_oracle_c.c"), "-lm"], check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
