"""CPU test of the field formation's sine and cosine (csrc/trig.hpp): the
same source compiled for the host, checked against FP64 sin/cos. Bars:
absolute error below 2.5e-7 for |x| <= 1e4 (dense grid, integer multiples
of pi/2 and pi/4 and their float neighbours), sin^2 + cos^2 = 1 within
5e-7, odd/even symmetry exact, and the quadrants at the floats nearest
k * pi/2."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mtd_transmission_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "trig.hpp"
int main(int argc, char** argv) {
  std::FILE* fi = std::fopen(argv[1], "rb");
  std::FILE* fo = std::fopen(argv[2], "wb");
  std::vector<float> x;
  float v;
  while (std::fread(&v, 4, 1, fi) == 1) x.push_back(v);
  for (float t : x) {
    float s, c;
    mtd::sincos_cw(t, &s, &c);
    std::fwrite(&s, 4, 1, fo);
    std::fwrite(&c, 4, 1, fo);
  }
  std::fclose(fi);
  std::fclose(fo);
  return 0;
}
"""


@pytest.fixture(scope="module")
def sincos(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("trig")
    src, exe = d / "drv.cpp", d / "drv"
    src.write_text(DRIVER)
    subprocess.run([HIPCC, "-O2", "-std=c++17", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True, capture_output=True)

    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        fi, fo = d / "in.bin", d / "out.bin"
        x.tofile(fi)
        subprocess.run([str(exe), str(fi), str(fo)], check=True)
        out = np.fromfile(fo, dtype=np.float32).reshape(-1, 2)
        return out[:, 0], out[:, 1]
    return run


def test_sincos_accuracy(sincos):
    rng = np.random.default_rng(7)
    k = np.arange(-6366, 6367, dtype=np.float64)  # multiples up to 1e4
    marks = np.concatenate([k * np.pi / 2, k * np.pi / 4]).astype(np.float32)
    near = np.concatenate([np.nextafter(marks, np.float32(np.inf)),
                           np.nextafter(marks, np.float32(-np.inf))])
    x = np.concatenate([np.linspace(-1e4, 1e4, 400001, dtype=np.float32),
                        rng.uniform(-8, 8, 200000).astype(np.float32),
                        marks, near, np.float32([0.0, -0.0, 1e-30, -1e-30])])
    s, c = sincos(x)
    xd = x.astype(np.float64)
    es = np.abs(s - np.sin(xd))
    ec = np.abs(c - np.cos(xd))
    assert es.max() < 2.5e-7, (es.max(), x[es.argmax()])
    assert ec.max() < 2.5e-7, (ec.max(), x[ec.argmax()])
    n = s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2
    assert np.abs(n - 1).max() < 5e-7
    # on |x| <= pi/4 (no reduction) the error is a few float ulps
    m = np.abs(xd) <= np.pi / 4
    assert es[m].max() < 1.2e-7 and ec[m].max() < 1.2e-7


def test_sincos_symmetry_and_quadrants(sincos):
    x = np.linspace(0, 1e3, 50001, dtype=np.float32)
    s, c = sincos(x)
    sn, cn = sincos(-x)
    assert np.array_equal(sn, -s)
    assert np.array_equal(cn, c)
    # the float nearest k pi/2 is off by up to 4e-7 from it: compare with the
    # FP64 values at the same floats, and the signs of the +-1 values
    q = (np.arange(-8, 9) * (np.pi / 2)).astype(np.float32)
    s, c = sincos(q)
    k = np.arange(-8, 9) % 4
    np.testing.assert_allclose(s, np.sin(q.astype(np.float64)), rtol=0, atol=1.2e-7)
    np.testing.assert_allclose(c, np.cos(q.astype(np.float64)), rtol=0, atol=1.2e-7)
    assert np.array_equal(np.rint(s), np.array([0, 1, 0, -1])[k])
    assert np.array_equal(np.rint(c), np.array([1, 0, -1, 0])[k])
