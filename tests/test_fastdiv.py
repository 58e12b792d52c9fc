"""The kernels' invariant-divisor division (svox2_amd/csrc/fastdiv.hpp)
against plain integer division, compiled for the host with g++: every
divisor 1..4096, powers of two and their neighbours up to 2^31, and grid
strides, each on numerators spread over [0, 2^31) including the values just
below and at multiples of the divisor."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(os.path.dirname(HERE), "svox2_amd", "csrc", "fastdiv.hpp")

PROG = r'''
#include "HDR"
#include <cstdio>
#include <vector>
using svox2amd::FastDiv;
static unsigned long long state = 88172645463325252ull;
static uint32_t rnd() { state ^= state << 13; state ^= state >> 7; state ^= state << 17;
                        return (uint32_t)(state >> 33); }  /* [0, 2^31) */
int main() {
    std::vector<uint32_t> ds;
    for (uint32_t d = 1; d <= 4096; ++d) ds.push_back(d);
    for (int k = 12; k <= 31; ++k) {
        uint32_t p = 1u << k;
        ds.push_back(p - 1); ds.push_back(p); if (k < 31) ds.push_back(p + 1);
    }
    const uint32_t strides[] = {65536u, 262144u, 16777216u, 134217728u, 255u * 255u,
                                511u * 511u, 17u * 12u, 300u * 2u, 70u * 65u};
    for (uint32_t s : strides) ds.push_back(s);
    long long checked = 0;
    for (uint32_t d : ds) {
        FastDiv f(d);
        for (int i = 0; i < 20000; ++i) {
            uint32_t n;
            if (i < 64) n = (uint32_t)i;
            else if (i < 128) n = 0x7fffffffu - (uint32_t)(i - 64);
            else if (i < 1000) { uint64_t q = rnd() / d; n = (uint32_t)(q * d + (i & 1 ? d - 1 : 0));
                                 if (n > 0x7fffffffu) n = 0x7fffffffu; }
            else n = rnd();
            uint32_t r;
            uint32_t q = f.divmod(n, r);
            if (q != n / d || r != n % d) { printf("FAIL d=%u n=%u q=%u\n", d, n, q); return 1; }
            ++checked;
        }
    }
    printf("ok %lld\n", checked);
    return 0;
}
'''


def test_fastdiv_matches_integer_division(tmp_path):
    src = tmp_path / "fastdiv_check.cpp"
    src.write_text(PROG.replace("HDR", HDR))
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["g++", "-O2", "-std=c++17", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
