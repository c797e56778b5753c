"""Host checks of the device math kernels in gwf_paths.h (exp10_kernel for the
interior grid points, sincos_kernel for the Earth-rotation angle of tau).

The same source is compiled for the host with g++ (IEEE double, fma, rint and
ldexp behave as on the GPU) and compared with long-double references on
random arguments over the ranges the kernels see, and wider.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

HARNESS = r'''
#include <cstdio>
#include <cmath>
#include <random>
#include "gwf_paths.h"
using namespace gwf;
int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-3.0, 6.0), V(-60.0, 60.0), W(-1e5, 1e5);
  const double eps = 2.220446049250313e-16;
  double me = 0, ms = 0, mc = 0;
  for (int t = 0; t < 400000; ++t) {
    const double x = U(rng);
    const long double ref = powl(10.0L, (long double)x);
    const double e = (double)(fabsl((exp10_kernel(x) - ref) / ref) / eps);
    me = e > me ? e : me;
    const double y = (t & 1) ? V(rng) : W(rng);
    double s, c;
    sincos_kernel(y, &s, &c);
    const double es = (double)(fabsl(s - sinl((long double)y)) / eps);
    const double ec = (double)(fabsl(c - cosl((long double)y)) / eps);
    ms = es > ms ? es : ms;
    mc = ec > mc ? ec : mc;
  }
  // quadrant and sign checks at multiples of pi/2 and at small arguments
  double s, c;
  sincos_kernel(0.0, &s, &c);
  const int ok0 = s == 0.0 && c == 1.0;
  sincos_kernel(-3.0 * 1.5707963267948966, &s, &c);
  const int okq = fabs(s - 1.0) < 4 * eps && fabs(c) < 4 * eps;
  sincos_kernel(1e-300, &s, &c);
  const int oks = s == 1e-300 && c == 1.0;
  printf("%.6f %.6f %.6f %d %d %d %.17g\n", me, ms, mc, ok0, okq, oks, exp10_kernel(2.0));
  return 0;
}
'''


@pytest.mark.skipif(shutil.which('g++') is None, reason='g++ not available')
def test_exp10_and_sincos_kernels(tmp_path):
    src = tmp_path / 'mathk.cpp'
    src.write_text(HARNESS)
    exe = tmp_path / 'mathk'
    subprocess.run(['g++', '-O2', '-std=c++17', '-I', os.path.join(ROOT, 'gwfast_amd', 'csrc'),
                    '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    exp_ulp, sin_eps, cos_eps = map(float, out[:3])
    ok0, okq, oks = map(int, out[3:6])
    assert exp_ulp <= 2.0          # relative error of 10^x, in ulp
    assert sin_eps <= 2.0 and cos_eps <= 2.0   # absolute error, in units of eps
    assert ok0 and okq and oks
    assert float(out[6]) == 100.0
