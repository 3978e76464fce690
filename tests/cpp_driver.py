"""Builds the C++ drivers under tests/cpp/ against include/ and libcola_spmm.so."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_driver(out_dir, name="drop_in_driver"):
    exe = os.path.join(str(out_dir), name)
    lib_dir = os.path.join(ROOT, "cola_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-o", exe,
           "-L", lib_dir, "-lcola_spmm", "-L", "/opt/rocm/lib", "-lamdhip64",
           f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.check_call(cmd)
    return exe
