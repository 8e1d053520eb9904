"""Build the C++ pipeline layer: libweavess_host.so and the weavess_main driver.

Host-only C++17 (g++), linked against libweavess_hip.so through the C ABI.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
SRCS = ["index.cpp", "components.cpp", "builder.cpp"]
LIB = os.path.join(HERE, "libweavess_host.so")
EXE = os.path.join(HERE, "weavess_main")
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-I", HERE]


def build(verbose: bool = False) -> None:
    link = ["-L", PKG, "-lweavess_hip", "-Wl,-rpath,$ORIGIN/.."]
    cmd = ["g++", *FLAGS, "-shared", "-o", LIB, *[os.path.join(HERE, s) for s in SRCS], *link]
    subprocess.run(cmd, check=True)
    cmd = ["g++", *FLAGS, "-o", EXE, os.path.join(HERE, "weavess_main.cpp"), "-L", HERE, "-lweavess_host",
           *link, "-Wl,-rpath,$ORIGIN"]
    subprocess.run(cmd, check=True)
    if verbose:
        print(f"built {LIB} and {EXE}")


if __name__ == "__main__":
    build(verbose=True)
