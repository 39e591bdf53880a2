"""Regenerates tests/golden/rng_libstdcxx.json from make_rng_fixture.cpp with
the image's g++ (libstdc++).  The fixture pins the oracle's restatement of
std::mt19937 and std::uniform_int_distribution<int>."""
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent

if __name__ == "__main__":
    with tempfile.TemporaryDirectory() as d:
        exe = Path(d) / "rngfix"
        subprocess.run(["g++", "-O2", "-std=c++17", str(HERE / "make_rng_fixture.cpp"), "-o", str(exe)],
                       check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    (HERE / "rng_libstdcxx.json").write_text(out)
    print("wrote", HERE / "rng_libstdcxx.json")
