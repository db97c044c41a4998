"""Builds the C++ host-layer test program into tests/cpp/_build/ (g++ against
katric_amd/_lib/libkatric_amd.so)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "test_cetric_api")


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.join(ROOT, "katric_amd", "_lib")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "test_cetric_api.cpp"), "-L", libdir, "-lkatric_amd",
                    "-Wl,-rpath," + libdir, "-o", OUT], check=True)
    return OUT


if __name__ == "__main__":
    print(build())
