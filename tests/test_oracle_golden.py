"""The oracle (CPU restatement, oracle/cpu_ref.hpp) against the reference
float build's golden vectors, and against what the mathematics fixes."""
import os

import numpy as np
import pytest

from tests import dumpfmt, golden_cases, oracle_lib


@pytest.mark.parametrize("name", golden_cases.names())
def test_oracle_matches_reference_every_step(name):
    g = golden_cases.load(name)
    w = oracle_lib.OracleWorld(g["desc_conf"])
    states = w.step_with_edits(g["steps"], g["edits"]) if len(g["edits"]) else w.step(g["steps"])
    bad = golden_cases.first_mismatch(g["digests"], states, dumpfmt)
    assert bad is None, "%s: first differing (step, part) = %s" % (name, bad)
    for st, full in g["full"].items():
        s = states[st - 1]
        np.testing.assert_array_equal(s["bodies"].view("<u4"), full["bodies"].view("<u4"))
        np.testing.assert_array_equal(s["ct"], full["ct"])


def test_hello_free_fall_matches_reference_expected_output():
    """HelloWorld/ExpectedOutput.txt (copied as data): the ball's x, y, angle
    at 2 decimals.  Lines 1-46 are free fall and must match; later lines differ
    because the reference program steps with TOI on (not in this restatement)."""
    g = golden_cases.load("hello")
    exp = [l.split() for l in open(os.path.join(golden_cases.GOLDEN, "hello_expected.txt"))]
    w = oracle_lib.OracleWorld(g["desc_conf"])
    states = w.step(46)
    for i, s in enumerate(states):
        x, y, a = s["bodies"][1][3:6]
        assert ["%.2f" % x, "%.2f" % y, "%.2f" % a] == exp[i], i


@pytest.mark.parametrize("precision", [32, 64])
def test_free_fall_closed_form(precision):
    """v_n = n a h and p_n = p_{n-1} + h v_n for a body with no contacts
    (the reference's GravitationalBodyMovement test, UnitTests/World.cpp:1492-1534)."""
    g = golden_cases.load("hello")
    w = oracle_lib.OracleWorld(g["desc_conf"], precision)
    dt = np.float32(1.0 / 60)  # the conf record carries dt as float
    a = np.float32(-9.8) if precision == 32 else np.float64(np.float32(-9.8))
    typ = np.float32 if precision == 32 else np.float64
    y = typ(4.0)
    v = typ(0.0)
    for n in range(1, 20):
        w.step(1, record=False)
        b = w.bodies_f64(2)[1]
        v = typ(v + typ(dt) * typ(a))
        y = typ(y + typ(dt) * v)
        assert b[4] == v and b[1] == y, (n, b, v, y)


def test_double_tracks_float_short_horizon():
    """The double instantiation agrees with the float one to float precision
    over a few pyramid steps (before the stack's chaotic divergence)."""
    g = golden_cases.load("pyramid")
    w32 = oracle_lib.OracleWorld(g["desc_conf"], 32)
    w64 = oracle_lib.OracleWorld(g["desc_conf"], 64)
    for _ in range(5):
        w32.step(1, record=False)
        w64.step(1, record=False)
    a = w32.bodies_f64(211)
    b = w64.bodies_f64(211)
    assert np.max(np.abs(a[:, :2] - b[:, :2])) < 1e-4


def test_host_sincos_matches_restated_routine():
    """The restated float sine/cosine used on the device (csrc/sincosf.hpp,
    compiled here for the host) equals the host C library on a dense sample;
    the full-range comparison over all 2^32 floats is done offline."""
    import ctypes
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = r'''
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include "sincosf.hpp"
int main() {
  long bad = 0;
  for (std::uint64_t b = 0; b < 0x7f800000ull; b += 4099) {
    for (int sg = 0; sg < 2; ++sg) {
      std::uint32_t u = std::uint32_t(b) | (sg ? 0x80000000u : 0u);
      float x; std::memcpy(&x, &u, 4);
      float c = std::cos(x), s = std::sin(x);
      float c1 = prx_math::CosF<true>(x), s1 = prx_math::SinF<true>(x);
      float s2, c2;
      prx_math::SinCosF<true>(x, &s2, &c2);
      bad += std::memcmp(&c, &c1, 4) != 0;
      bad += std::memcmp(&s, &s1, 4) != 0;
      bad += std::memcmp(&c, &c2, 4) != 0;
      bad += std::memcmp(&s, &s2, 4) != 0;
    }
  }
  std::printf("%ld\n", bad);
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        open(cpp, "w").write(src)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(root, "playrho_amd", "csrc"), cpp, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == 0
