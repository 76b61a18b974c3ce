"""The C++ facade (include/cdft/dft.hpp) on the CPU: it compiles against the
C ABI with g++, and its argument checks raise the reference's exception types
with the reference's messages (grid.hpp:87-103, models.hpp:398-410) before any
device work.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def _binary():
    subprocess.run(["make", "-s", "-C", CPP], check=True)
    return os.path.join(CPP, "facade_check")


def test_facade_compiles_and_maps_errors():
    r = subprocess.run([_binary(), "cpu"], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade cpu checks: ok" in r.stdout
