"""The aggregation kernel's mean sum / cnt (cnt = 1 .. 121) is computed as
q = sum * r, q' = fma(fma(-q, cnt, sum), r, q) with r = 1/cnt rounded to
nearest (Markstein's correction). This checks, on the CPU's IEEE float32
arithmetic with fused multiply-add, that q' equals the correctly rounded
quotient: exhaustively for every float in [1, 2) and every divisor (the
result for a * 2^k follows by exact scaling while the quotient stays
normal), and on random floats of every binade, where the only mismatches
allowed are subnormal quotients, off by one subnormal ulp."""
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float mean(float a, int c, const float* rc) {
  const float r = rc[c], q = a * r;
  return fmaf(fmaf(-q, (float)c, a), r, q);
}
int main(void) {
  float rc[128];
  for (int c = 1; c < 128; c++) rc[c] = 1.0f / (float)c;
  long bad_normal = 0, bad_sub = 0, worst_sub = 0;
  for (int c = 1; c <= 121; c++)
    for (uint32_t u = 0x3f800000u; u < 0x40000000u; u++) {
      const float a = f_of(u);
      if (u_of(mean(a, c, rc)) != u_of(a / (float)c)) bad_normal++;
    }
  uint64_t s = 88172645463325252ull;
  for (long i = 0; i < 30000000; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    const float a = f_of((uint32_t)(s % 0x7f800000u));
    const int c = 1 + (int)((s >> 40) % 121);
    const float got = mean(a, c, rc), ref = a / (float)c;
    if (u_of(got) == u_of(ref)) continue;
    if (fabsf(ref) >= 0x1p-126f) { bad_normal++; continue; }
    bad_sub++;
    long d = labs((long)u_of(got) - (long)u_of(ref));
    if (d > worst_sub) worst_sub = d;
  }
  printf("%ld %ld %ld\n", bad_normal, bad_sub, worst_sub);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_markstein_mean_is_correctly_rounded(tmp_path):
    src = tmp_path / "m.c"
    exe = tmp_path / "m"
    src.write_text(SRC)
    # -ffp-contract=off: a * r and the residual stay separately rounded
    subprocess.run(["gcc", "-O2", "-march=native", "-ffp-contract=off", str(src), "-o", str(exe),
                    "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=600)
    bad_normal, bad_sub, worst_sub = map(int, out.stdout.split())
    print(f"normal-quotient mismatches {bad_normal}, subnormal {bad_sub} (worst {worst_sub} ulp)")
    assert bad_normal == 0
    assert worst_sub <= 1
