"""C++ checks of the host library: tests/cpp/host_checks.cpp (the threaded
name audit of Model::validate reports the first offending index as a serial
scan would, and verify_primal's threaded long-double row activities equal
SparseMatrixCSC::multiply bit for bit, on models past the threading
threshold) and tests/cpp/pack_checks.cpp (the batch packing the GPU solve
loads: 16 threads pack exactly what one packs, and every packed entry, bound
and objective coefficient is the original one through the relabelling)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "markov-zip1_amd", "host")
LIB = os.path.join(ROOT, "markov-zip1_amd", "lib")


def run_check(tmp_path, name):
    exe = tmp_path / name
    src = os.path.join(ROOT, "tests", "cpp", f"{name}.cpp")
    b = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-I", HOST, src, "-o", str(exe), "-L", LIB,
                        "-lmarkov_cero_pdlp", f"-Wl,-rpath,{LIB}"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr


def test_host_checks(tmp_path):
    run_check(tmp_path, "host_checks")


def test_pack_checks(tmp_path):
    run_check(tmp_path, "pack_checks")
