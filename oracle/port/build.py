"""Build the oracle's single-thread C++ restatement of the reference's
1-qubit step (cpu_port.cpp) into oracle/_build/libbass_cpu.so with g++ -O3.
Oracle code: loaded only by the tests and the benchmark's CPU baseline."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(os.path.dirname(HERE), "_build")
OUT = os.path.join(OUT_DIR, "libbass_cpu.so")


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    # x86-64-v3 (AVX2 + FMA): the hosts this runs on; std::fma maps to one instruction
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-shared", "-fPIC",
           os.path.join(HERE, "cpu_port.cpp"), "-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(OUT + ".tmp", OUT)
    print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
