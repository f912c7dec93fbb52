"""The float64 erf of the HIP kernels (erf_tab, xpg_amd/csrc/xpg_erf.h) and its table (xpg_erf_table.h).

Pins: the table entries are erf and erf' at the nodes, correctly rounded, checked against 80-digit decimal values,
scipy and Python's libm erf; the header compiled for the host gives, bit for bit, the operations that
gen_erf_table.py emulates with exact rational arithmetic; that emulation is within 1.3 ulp of 80-digit erf; odd
symmetry, saturation beyond 6 and NaN propagation.  The erfc column and erfc_tab likewise: entries against 80-digit
values and scipy, the host build against the emulation, the emulation within 2e-13 (relative) of 80-digit erfc; and
the oracle's erf_diff against 80-digit differences in the tails, where erf(a) - erf(b) itself would cancel.
"""
import math
import random
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
from scipy.special import erf

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "xpg_amd" / "csrc"
sys.path.insert(0, str(CSRC))
import gen_erf_table as G   # noqa: E402

HIPCC = Path("/opt/rocm/bin/hipcc")


def _header_table():
    text = (CSRC / "xpg_erf_table.h").read_text()
    body = text[text.index("ERF_TAB["):text.index("ERFC_TAB[")]
    pairs = re.findall(r"\{([-0-9.e+]+), ([-0-9.e+]+)\}", body)
    return [(float(a), float(b)) for a, b in pairs]


def _header_erfc():
    text = (CSRC / "xpg_erf_table.h").read_text()
    pairs = re.findall(r"\{([-0-9.e+]+), ([-0-9.e+]+)\}", text[text.index("ERFC_TAB["):])
    tab = _header_table()
    assert [float(b) for _, b in pairs] == [t[1] for t in tab]        # the erf' column, repeated
    return [float(a) for a, _ in pairs]


@pytest.fixture(scope="module")
def host_erf(tmp_path_factory):
    if not HIPCC.exists():
        pytest.skip("hipcc is not installed")
    exe = tmp_path_factory.mktemp("erf") / "erf_host"
    subprocess.run([str(HIPCC), "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950", f"-I{CSRC}",
                    f"-I{ROOT / 'include'}", str(ROOT / "tests" / "erf_host.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)

    def run(xs, col=0):
        out = subprocess.run([str(exe)], input="\n".join(repr(float(x)) for x in xs), capture_output=True, text=True,
                             check=True).stdout.split()
        return [float(v) for v in out[col::2]]
    return run


def test_table_entries_are_correctly_rounded():
    tab = _header_table()
    assert len(tab) == G.N == 6 * 256 + 1
    x = np.arange(G.N) / 256.0
    f0 = np.array([t[0] for t in tab])
    f1 = np.array([t[1] for t in tab])
    # independent implementations (each good to a unit or two in the last place) agree everywhere
    assert np.all(np.abs(f0 - erf(x)) <= 2 * np.spacing(np.abs(f0)))
    assert np.all(np.abs(f0 - np.array([math.erf(v) for v in x])) <= np.spacing(np.abs(f0)))
    d1 = 2.0 / math.sqrt(math.pi) * np.exp(-x * x)
    assert np.all(np.abs(f1 - d1) <= 3 * np.spacing(np.abs(f1)))
    # the 80-digit values, correctly rounded, at a sample of nodes (first, last and every 61st)
    for i in list(range(0, G.N, 61)) + [G.N - 1]:
        xd = G.D(i) / 256
        assert f0[i] == float(G.erf_dec(xd)), i
        assert f1[i] == float(G.TWO_OVER_SQRTPI * (-(xd * xd)).exp()), i
    assert f0[0] == 0.0 and f0[-1] == 1.0


def test_host_build_equals_exact_emulation(host_erf):
    tab = _header_table()
    rng = random.Random(17)
    xs = [rng.uniform(-7.0, 7.0) for _ in range(4000)] + [rng.uniform(-0.01, 0.01) for _ in range(300)]
    xs += [i / 256 + s / 512 for i in range(0, G.N - 1, 29) for s in (-1, 1)]          # halfway between nodes
    xs += [0.0, -0.0, 1e-300, -1e-300, 6.0, -6.0, 6.0 + 2 ** -50, 5.93, 1e300, math.inf, -math.inf]
    got = host_erf(xs)
    want = [G.erf_model(tab, x) for x in xs]
    bad = [(x, g, w) for x, g, w in zip(xs, got, want) if not (g == w and math.copysign(1, g) == math.copysign(1, w))]
    assert not bad, bad[:5]
    assert math.isnan(host_erf([math.nan])[0])


def test_emulation_error_against_80_digit_erf():
    tab = _header_table()
    rng = random.Random(23)
    xs = [rng.uniform(0.0, 6.5) for _ in range(600)] + [rng.uniform(0.0, 0.02) for _ in range(100)]
    worst = 0.0
    for x in xs:
        ref = G.erf_dec(x)
        got = G.erf_model(tab, x)
        err = abs(G.D(got) - ref)
        if float(ref):
            worst = max(worst, float(err) / math.ulp(float(ref)))
        assert G.erf_model(tab, -x) == -got
    assert worst <= 1.3, worst
    assert all(G.erf_model(tab, x) == 1.0 for x in (6.0 + 1e-9, 7.0, 40.0, math.inf))


def test_erfc_entries_are_correctly_rounded():
    from scipy.special import erfc
    c = np.array(_header_erfc())
    assert c.size == G.N
    x = np.arange(G.N) / 256.0
    assert np.all(np.abs(c - erfc(x)) <= 8 * np.spacing(c))          # scipy erfc: within a few ulp
    for i in list(range(0, G.N, 97)) + [G.N - 1]:
        assert c[i] == float(1 - G.erf_dec(G.D(i) / 256)), i
    assert c[0] == 1.0 and 2.1e-17 < c[-1] < 2.2e-17


def test_erfc_host_build_equals_exact_emulation(host_erf):
    tab, ctab = _header_table(), _header_erfc()
    rng = random.Random(19)
    xs = [rng.uniform(0.0, 6.0) for _ in range(3000)] + [i / 256 + s / 512 for i in range(1, G.N - 1, 31)
                                                        for s in (-1, 1)] + [0.0, 6.0, 5.999999999]
    got = host_erf(xs, col=1)
    want = [G.erfc_model(tab, ctab, x) for x in xs]
    bad = [(x, g, w) for x, g, w in zip(xs, got, want) if g != w]
    assert not bad, bad[:5]
    assert host_erf([6.0 + 2 ** -40, 6.5, 40.0, math.inf], col=1) == [0.0] * 4     # beyond the table
    assert math.isnan(host_erf([math.nan], col=1)[0])


def test_erfc_emulation_relative_error_against_80_digit_erfc():
    tab, ctab = _header_table(), _header_erfc()
    rng = random.Random(29)
    worst = 0.0
    for x in [rng.uniform(0.0, 6.0) for _ in range(500)] + [5.99, 5.5, 4.0]:
        ref = 1 - G.erf_dec(x)
        worst = max(worst, float(abs(G.D(G.erfc_model(tab, ctab, x)) - ref) / ref))
    assert worst <= 2e-13, worst


def test_oracle_erf_diff_keeps_tail_precision():
    """The oracle's rocking-curve erf differences: tail pairs of one sign against 80-digit differences (relative
    1e-13; the plain difference of erf values is off by up to 1e-4 there), pairs across zero against erf(a) - erf(b)."""
    sys.path.insert(0, str(ROOT))
    from oracle.kernels import erf_diff
    for a, b in [(5.2, 4.9), (4.0, 3.7), (-3.7, -4.0), (5.9, 5.7), (2.0, 1.9)]:
        ref = G.erf_dec(a) - G.erf_dec(b)
        got = float(erf_diff(a, b))
        assert abs(G.D(got) - ref) / abs(ref) < 1e-13, (a, b, got, float(ref))
    assert float(erf_diff(6.3, 6.1)) == 0.0 == float(erf(6.3) - erf(6.1))     # beyond 6: 0, as erf differences
    plain = float(erf(5.2) - erf(4.9))
    assert abs(G.D(plain) - (G.erf_dec(5.2) - G.erf_dec(4.9))) / (G.erf_dec(5.2) - G.erf_dec(4.9)) > 1e-9
    assert abs(float(erf_diff(0.3, -0.4)) - float(erf(0.3) - erf(-0.4))) <= 4e-16
