"""The re-sort's space-filling-curve keys (yukawamd_amd/csrc/yk_sfc.hpp),
compiled for the host with g++: on every grid from 2 to 128 cells per axis
the keys run over [0, 8^g) once each, and consecutive keys are face-adjacent
cells (the Hilbert property that keeps 16- and 64-slot groups compact; a
Morton order fails it)."""
import shutil
import subprocess
from pathlib import Path

import pytest

HDR = Path(__file__).resolve().parent.parent / "yukawamd_amd" / "csrc"

PROGRAM = r"""
#define __host__
#define __device__
#include "yk_sfc.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>
int main() {
    for (int g = 1; g <= 7; ++g) {
        const int G = 1 << g;
        std::vector<int> cell((size_t)G * G * G, -1);
        for (int x = 0; x < G; ++x)
            for (int y = 0; y < G; ++y)
                for (int z = 0; z < G; ++z) {
                    const unsigned k = yk::hilbert_key3(x, y, z, g);
                    if (k >= cell.size() || cell[k] >= 0) { std::printf("not a bijection g=%d\n", g); return 1; }
                    cell[k] = (x * G + y) * G + z;
                }
        for (size_t k = 1; k < cell.size(); ++k) {
            const int a = cell[k - 1], b = cell[k];
            const int d = std::abs(a / (G * G) - b / (G * G)) + std::abs(a / G % G - b / G % G) +
                          std::abs(a % G - b % G);
            if (d != 1) { std::printf("keys %zu, %zu not adjacent g=%d\n", k - 1, k, g); return 1; }
        }
    }
    std::printf("ok\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_hilbert_keys_bijective_and_adjacent(tmp_path):
    src = tmp_path / "sfc.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "sfc"
    subprocess.run(["g++", "-O2", "-I", str(HDR), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
