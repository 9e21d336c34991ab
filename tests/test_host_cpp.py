"""C++ checks of the host library (tests/cpp/host_checks.cpp): the threaded
name audit of Model::validate reports the first offending index as a serial
scan would, and verify_primal's threaded long-double row activities equal
SparseMatrixCSC::multiply bit for bit, on models past the threading threshold."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "markov-zip1_amd", "host")
LIB = os.path.join(ROOT, "markov-zip1_amd", "lib")


def test_host_checks(tmp_path):
    exe = tmp_path / "host_checks"
    src = os.path.join(ROOT, "tests", "cpp", "host_checks.cpp")
    b = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-I", HOST, src, "-o", str(exe), "-L", LIB,
                        "-lmarkov_cero_pdlp", f"-Wl,-rpath,{LIB}"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
