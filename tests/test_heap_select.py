"""The product's heap-select emulation (faust_amd/csrc/heap_select.hpp, used by
the projector kernels) selects exactly the set std::partial_sort(|a|>|b|)
keeps, ties included: compiled for the CPU with g++ and run on tie-heavy data."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "heap_select.hpp"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <set>
#include <vector>
int main() {
  std::mt19937 g(12345);
  long bad = 0, tot = 0;
  for (int trial = 0; trial < 100000; ++trial) {
    int n = 2 + g() % 64;
    int k = 1 + g() % (n - 1);
    int nv = 1 + g() % 5;
    std::vector<double> v(n);
    for (auto& x : v) x = (double)(int)(g() % nv) * ((g() & 1) ? 1 : -1) * 0.25;
    std::vector<std::pair<int, double>> vp(n);
    for (int i = 0; i < n; ++i) vp[i] = {i, v[i]};
    std::partial_sort(vp.begin(), vp.begin() + k, vp.end(),
                      [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                        return std::fabs(a.second) > std::fabs(b.second);
                      });
    std::set<int> ref;
    for (int i = 0; i < k; ++i) ref.insert(vp[i].first);
    std::vector<int> hi(k);
    std::vector<double> ha(k);
    fh::heap_select<double>(n, k, [&](int i) { return std::fabs(v[i]); }, hi.data(), ha.data());
    std::set<int> got(hi.begin(), hi.end());
    ++tot;
    if (got != ref) ++bad;
  }
  std::printf("%ld %ld\n", bad, tot);
  return bad != 0;
}
'''


def test_heap_select_matches_partial_sort(tmp_path):
    src = tmp_path / "hs.cpp"
    src.write_text(SRC)
    exe = tmp_path / "hs"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "faust_amd", "csrc"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    bad, tot = map(int, out.stdout.split())
    assert tot == 100000 and bad == 0
