"""lattice_skip (monkeygl_amd/csrc/mk_lattice.h), the ray casters' walk over
empty cells, against the loop it replaces:

    while (acc + l1 < a_skip && it < max_iter) { acc += l1; ++it; }

The VR and MIP kernels skip the lattice points of an empty cell with it
(core/kernel.cu:133-264 visits every one of them; a point in an empty cell
changes nothing but acc), so its floats must be exactly the loop's. The header
is compiled for the host with g++ and run on seeded random cases, on binade
boundaries, on steps that fall exactly halfway between two floats, on steps
too small to move acc, and on negative, tiny and zero starts."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HDR = ROOT / "monkeygl_amd" / "csrc" / "mk_lattice.h"

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <random>
#include "mk_lattice.h"

static void loop_ref(float& acc, float l1, float a_skip, int& it, int max_iter) {
    while (acc + l1 < a_skip && it < max_iter) { acc += l1; ++it; }
}
static long n_cases = 0, n_bad = 0, n_moved = 0;
static void check(float acc, float l1, float a_skip, int it, int max_iter) {
    float a0 = acc, a1 = acc; int i0 = it, i1 = it;
    loop_ref(a0, l1, a_skip, i0, max_iter);
    mkr::lattice_skip(a1, l1, 1.0f / l1, a_skip, i1, max_iter);
    ++n_cases;
    if (i0 != it) ++n_moved;
    if (std::memcmp(&a0, &a1, 4) != 0 || i0 != i1) {
        if (n_bad < 10)
            std::printf("MISMATCH acc=%a l1=%a a_skip=%a it=%d max=%d: loop %a/%d skip %a/%d\n",
                        acc, l1, a_skip, it, max_iter, a0, i0, a1, i1);
        ++n_bad;
    }
}
int main() {
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    // ray parameters of the renderers: acc in [0, 1.8], steps 1/(n * k)
    for (int c = 0; c < 400000; ++c) {
        const double n = 8 + U(g) * 2000;
        const float l1 = (float)(1.0 / (n * (1.0 + 3.0 * U(g))));
        float acc = (float)(U(g) * 1.8);
        if (c % 3 == 0) {  // a point of the accumulation from 0
            int k = (int)(U(g) * 600);
            acc = 0.f;
            for (int i = 0; i < k; ++i) acc += l1;
        }
        const float span = (float)(U(g) * 300.0) * l1;
        const float a_skip = acc + span;
        const int it = (int)(U(g) * 50);
        const int max_iter = (c % 5 == 0) ? it + (int)(U(g) * 40) : 1 << 20;
        check(acc, l1, a_skip, it, max_iter);
    }
    // bounds exactly on points of the accumulation
    for (int c = 0; c < 20000; ++c) {
        const float l1 = (float)(1.0 / (16 + U(g) * 3000));
        float acc = (float)(U(g) * 1.5);
        float b = acc;
        const int k = 1 + (int)(U(g) * 200);
        for (int i = 0; i < k; ++i) b += l1;
        check(acc, l1, b, 0, 1 << 20);
        check(acc, l1, std::nextafter(b, 2.0f), 0, 1 << 20);
        check(acc, l1, std::nextafter(b, 0.0f), 0, 1 << 20);
    }
    // walks across binade boundaries (0.25, 0.5, 1, 2)
    for (int c = 0; c < 40000; ++c) {
        const float edge = std::ldexp(1.0f, -2 + (int)(U(g) * 4));
        const float l1 = (float)(edge * (1e-5 + U(g) * 0.02));
        const float acc = edge - (float)(U(g) * 40) * l1;
        check(acc, l1, edge + (float)(U(g) * 80) * l1, 0, 1 << 20);
    }
    // steps halfway between two floats of acc's binade (ties), and steps that
    // are an exact number of floats
    for (int c = 0; c < 20000; ++c) {
        const int e = -3 + (int)(U(g) * 4);
        const float u = std::ldexp(1.0f, e - 23);
        const float acc = std::ldexp(1.0f, e) * (float)(1.0 + U(g) * 0.9);
        const float accq = std::ldexp(std::floor(std::ldexp(acc, 23 - e)), e - 23);
        const int D = 1 + (int)(U(g) * 5000);
        const float l1 = (c & 1) ? (float)((D + 0.5) * u) : (float)(D * (double)u);
        check(accq, l1, accq + (float)(U(g) * 500) * l1, 0, 1 << 20);
    }
    // steps of a few floats over long spans (jumps of up to ~2^23 additions)
    for (int c = 0; c < 300; ++c) {
        const int e = -2 + (int)(U(g) * 3);
        const float u = std::ldexp(1.0f, e - 23);
        const float acc = std::ldexp(1.0f, e) * (float)(1.0 + U(g) * 0.05);
        const float l1 = (float)((1 + (int)(U(g) * 12) + U(g)) * (double)u);
        check(acc, l1, acc * (float)(1.0 + U(g) * 1.5), 0, 1 << 30);
    }
    // steps of 8 to 300 floats, where the count's estimate is loosest, over
    // spans around the jump cap (a quarter of the step in floats)
    for (int c = 0; c < 200000; ++c) {
        const int e = -3 + (int)(U(g) * 4);
        const float u = std::ldexp(1.0f, e - 23);
        const float acc = std::ldexp(1.0f, e) * (float)(1.0 + U(g) * 0.999);
        const double qd = 8.0 + U(g) * 292.0;
        const float l1 = (float)(qd * (double)u);
        const float span = (float)((0.5 + U(g) * 0.5) * (qd / 4 + 3)) * l1;
        check(acc, l1, acc + span, 0, 1 << 20);
    }
    // steps too small to move acc (the loop counts out max_iter), zero,
    // negative and tiny starts, huge and non-finite bounds
    for (int c = 0; c < 2000; ++c) {
        const float acc = (float)(0.5 + U(g));
        const float l1 = std::ldexp(acc, -26) * (float)U(g);
        check(acc, l1, acc * 2.0f, (int)(U(g) * 10), 10 + (int)(U(g) * 3000));
    }
    for (int c = 0; c < 20000; ++c) {
        const float l1 = (float)(1e-4 + U(g) * 1e-2);
        const float starts[] = {0.0f, -0.0f, -(float)U(g) * 0.01f, 1e-38f, 1e-31f, std::ldexp(1.0f, -100),
                                std::ldexp(1.0f, -101), (float)U(g) * 1e-6f};
        for (float a : starts) {
            check(a, l1, a + (float)(U(g) * 200) * l1, 0, 1 << 20);
            check(a, l1, 1e30f, 0, 5000);
            check(a, l1, INFINITY, 3, 4000);
            check(a, l1, NAN, 0, 100);
        }
    }
    std::printf("cases %ld moved %ld bad %ld\n", n_cases, n_moved, n_bad);
    return n_bad != 0;
}
"""


def _compile(tmp_path: Path) -> Path:
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "lattice.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "lattice"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-D__host__=", "-D__device__=",
                    "-I", str(HDR.parent), str(src), "-o", str(exe)], check=True)
    return exe


def test_lattice_skip_equals_the_loop(tmp_path):
    exe = _compile(tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    cases, moved, bad = (int(w) for w in out.stdout.split()[1::2])
    assert bad == 0 and cases > 500000 and moved > cases // 2


def test_lattice_skip_check_catches_a_one_point_slip(tmp_path):
    # the harness against a loop with a plausible mistake planted in it (<=
    # for <, one point too many whenever the bound is itself a point of the
    # accumulation) must report mismatches
    bad = HARNESS.replace("while (acc + l1 < a_skip && it < max_iter)", "while (acc + l1 <= a_skip && it < max_iter)")
    assert bad != HARNESS
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    cpp = tmp_path / "lattice_bad.cpp"
    cpp.write_text(bad)
    exe = tmp_path / "lattice_bad"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-D__host__=", "-D__device__=",
                    "-I", str(HDR.parent), str(cpp), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode != 0 and "MISMATCH" in out.stdout
