"""The compiled mixed-radix plans (qstem_amd/csrc/qs_mix.hpp) on the CPU: every plan
of the MixPlan table is a valid factorisation for the kernels' lane layout, its LDS
slot holds every exchange position, and the kernels' Stockham index scheme (butterfly
j = t + T b of lane t, inputs j + k N/R, twiddles W_N^(k (j mod NS) N/(NS R)), outputs
(j / NS) NS R + j mod NS + k NS, element t + T m in register m) run with numpy DFTs
gives numpy's FFT in both directions.  The lengths listed in QS_MIX_SIZES (which the
build compiles) are exactly the table's."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "qstem_amd", "csrc")


def _plans():
    text = open(os.path.join(CSRC, "qs_mix.hpp")).read()
    out = {}
    for m in re.finditer(r"struct MixPlan<(\d+)> \{ static constexpr int ([^;]+);", text):
        fields = dict((k.strip(), int(v)) for k, v in (f.split("=") for f in m.group(2).split(",")))
        out[int(m.group(1))] = fields
    return out


def _sizes():
    text = open(os.path.join(CSRC, "qs_internal.hpp")).read()
    m = re.search(r"#define QS_MIX_SIZES\(X\)((?:[ \t]*\\\n|[ \t]*X\(\d+\))+)", text)
    return [int(n) for n in re.findall(r"X\((\d+)\)", m.group(1))]


PLANS = _plans()


def _pos(p, n, i):
    if p["NP"] == 2:
        return (i % p["R0"]) * (n // p["R0"] + 1) + i // p["R0"]
    return i + (i >> 5)


def _transform(p, n, x, sign):
    e = p["E"]
    t_n = n // e
    radices = [p["R0"], p["R1"], p["R2"], p["R3"]][:p["NP"]]
    v = np.array([[x[t + t_n * m] for m in range(e)] for t in range(t_n)], dtype=complex)
    slot = np.full(p["LD"], np.nan + 0j)
    ns = 1
    for ip, r in enumerate(radices):
        nb = e // r
        first, last = ip == 0, ip == len(radices) - 1
        dft = np.exp(sign * 2j * np.pi * np.outer(np.arange(r), np.arange(r)) / r)
        outs = []
        for t in range(t_n):
            for b in range(nb):
                j = t + t_n * b
                a = np.array([v[t][b + k * nb] if first else slot[_pos(p, n, j + k * (n // r))] for k in range(r)])
                if ns > 1:
                    a = a * np.exp(sign * 2j * np.pi * np.arange(r) * ((j % ns) * (n // (ns * r))) / n)
                outs.append((t, b, j, dft @ a))
        if last:
            for t, b, j, y in outs:
                for k in range(r):
                    v[t][b + k * nb] = y[k]
        else:
            slot[:] = np.nan  # a position read before it is written shows up as NaN
            for t, b, j, y in outs:
                base = (j // ns) * ns * r + j % ns
                for k in range(r):
                    slot[_pos(p, n, base + k * ns)] = y[k]
        ns *= r
    return np.array([v[i % t_n][i // t_n] for i in range(n)])


def test_table_matches_the_compiled_list():
    assert sorted(PLANS) == sorted(_sizes())


@pytest.mark.parametrize("n", sorted(PLANS))
def test_plan_layout(n):
    p = PLANS[n]
    radices = [p["R0"], p["R1"], p["R2"], p["R3"]][:p["NP"]]
    assert np.prod(radices) == n and p["NP"] in (2, 3, 4)
    assert all(r == 1 for r in [p["R0"], p["R1"], p["R2"], p["R3"]][p["NP"]:])
    assert all(p["E"] % r == 0 for r in radices)
    t_n = n // p["E"]
    assert n % p["E"] == 0 and t_n <= 64 and p["E"] <= 32
    assert n // radices[-1] == t_n * (p["E"] // radices[-1])  # the last pass ends in the natural layout
    assert max(_pos(p, n, i) for i in range(n)) < p["LD"] and n <= p["LD"]
    assert len({_pos(p, n, i) for i in range(n)}) == n


@pytest.mark.parametrize("n", sorted(PLANS))
def test_plan_index_scheme_is_the_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    p = PLANS[n]
    fwd = _transform(p, n, x, -1)
    bwd = _transform(p, n, x, +1)
    assert np.max(np.abs(fwd - np.fft.fft(x))) < 1e-9 * n
    assert np.max(np.abs(bwd - n * np.fft.ifft(x))) < 1e-9 * n
