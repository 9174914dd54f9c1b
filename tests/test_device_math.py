"""The table-driven e^{-a}, sin and cos of the kernels (cqbem_device.hpp: exp_neg_tab, sincos_tab), built for
the host from the same header and checked against mpmath at 200 bits:
- e^{-a} within 0.52 ulp of the result, a in [0, 740) (log-uniform below 1, uniform above), and the clamps;
- sin y and cos y within 0.3 ulp of 1, |y| up to 1e6 (the reduction's supported range), 2 and 1e-3;
- exact values at 0: e^0 = 1, sin 0 = 0, cos 0 = 1."""
import math
import os
import random
import subprocess

import pytest

mp = pytest.importorskip("mpmath")

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def check_bin(tmp_path_factory):
    out = tmp_path_factory.mktemp("devmath") / "check"
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    r = subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-o", str(out),
                        os.path.join(HERE, "device_math_check.cpp")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return str(out)


def run(binary, lines):
    r = subprocess.run([binary], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [[float.fromhex(t) for t in l.split()] for l in r.stdout.strip().split("\n")]


def test_exp(check_bin):
    mp.mp.prec = 200
    rng = random.Random(1)
    args = [math.exp(rng.uniform(-30, 0)) for _ in range(1500)] + [rng.uniform(0, 740) for _ in range(1500)]
    args += [0.0, 1e-300, 700.0, 708.0]
    got = run(check_bin, [f"e {a!r}" for a in args])
    worst = 0.0
    for a, (v,) in zip(args, got):
        ref = mp.exp(-mp.mpf(a))
        if ref < mp.mpf(2) ** -1022:  # subnormal results: absolute error below the smallest normal's ulp
            assert abs(mp.mpf(v) - ref) <= mp.mpf(2) ** -1074
            continue
        worst = max(worst, float(abs(mp.mpf(v) - ref)) / math.ulp(float(ref)))
    assert worst <= 0.52, worst
    assert run(check_bin, ["e 0.0"])[0][0] == 1.0
    # the clamps: e^{-a} for a > 740 is e^{-740}, for a < -700 e^{700}
    big, neg = run(check_bin, ["e 800.0", "e -800.0"])
    assert big[0] == run(check_bin, ["e 740.0"])[0][0] and neg[0] == run(check_bin, ["e -700.0"])[0][0]


@pytest.mark.parametrize("span", [1e6, 600.0, 2.0, 1e-3])
def test_sincos(check_bin, span):
    mp.mp.prec = 200
    rng = random.Random(int(span * 7) + 3)
    ys = [rng.uniform(-span, span) for _ in range(2000)]
    got = run(check_bin, [f"t {y!r}" for y in ys])
    ws = wc = 0.0
    for y, (s, c) in zip(ys, got):
        ws = max(ws, float(abs(mp.mpf(s) - mp.sin(mp.mpf(y)))) / math.ulp(1.0))
        wc = max(wc, float(abs(mp.mpf(c) - mp.cos(mp.mpf(y)))) / math.ulp(1.0))
    assert ws <= 0.3 and wc <= 0.3, (ws, wc)
    assert run(check_bin, ["t 0.0"])[0] == [0.0, 1.0]
