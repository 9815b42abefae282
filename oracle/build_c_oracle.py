"""Build recipe of the C restatement (oracle/cpu_md.c) -- test infrastructure.

Two variants are built so the prebuilt library runs on any x86-64 host:
x86-64-v3 (AVX2/FMA units, used when the CPU has them) and x86-64-v2.
-ffp-contract=off keeps every float operation rounded once, in source order.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_md.c")
OUTS = {"v3": os.path.join(HERE, "libcpu_md_v3.so"), "v2": os.path.join(HERE, "libcpu_md_v2.so")}


def build(verbose=False):
    for level, out in OUTS.items():
        cmd = ["gcc", "-O3", f"-march=x86-64-{level}", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
               "-fPIC", "-shared", SRC, "-o", out + ".tmp", "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
    return OUTS


def library_path():
    flags = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    flags = line
                    break
    except OSError:
        pass
    if " avx2" in flags and " fma" in flags and os.path.exists(OUTS["v3"]):
        return OUTS["v3"]
    return OUTS["v2"]


if __name__ == "__main__":
    print(build(verbose=True))
