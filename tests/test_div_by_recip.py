"""The pixel-centre quotient (x + 0.5) / w by a shared reciprocal
(div_by_recip in violet_amd/csrc/vgi_common.h, Markstein's division) equals
the IEEE division for every x < w <= 16384: the host build of the same
header function, checked exhaustively (134M quotients), and a planted
error (the quotient without its correction step) is caught."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

SRC = r"""
#include <cmath>
#include <cstdio>
#include <cstring>
#include "vgi_common.h"
int main() {
    unsigned long long bad = 0, n = 0;
    for (uint32_t w = 1; w <= 16384; ++w) {
        volatile float fw = (float)w;
        const float y = 1.0f / fw;
        for (uint32_t c = 0; c < w; ++c) {
            const float a = (float)c + 0.5f;
            volatile float ref = a / fw;
#ifdef PLANT
            const float got = a * y;
#else
            const float got = vgi::div_by_recip(a, fw, y);
#endif
            float r = ref;
            if (std::memcmp(&got, &r, 4) != 0) ++bad;
            ++n;
        }
    }
    std::printf("%llu %llu\n", n, bad);
    return 0;
}
"""


def _run(tmp_path: Path, plant: bool) -> tuple[int, int]:
    src = tmp_path / "div.cpp"
    src.write_text(SRC)
    exe = tmp_path / ("div_plant" if plant else "div")
    cmd = ["g++", "-std=c++17", "-O2", "-mfma", "-ffp-contract=off", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'violet_amd' / 'csrc'}", str(src), "-o", str(exe)] + (["-DPLANT"] if plant else [])
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=300).stdout.split()
    return int(out[0]), int(out[1])


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pixel_centre_quotients_equal_ieee_division(tmp_path):
    n, bad = _run(tmp_path, plant=False)
    assert n == 16384 * 16385 // 2
    assert bad == 0


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_uncorrected_quotient_is_caught(tmp_path):
    n, bad = _run(tmp_path, plant=True)
    assert bad > 0
