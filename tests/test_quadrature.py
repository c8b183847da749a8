"""The literal t-quadrature tables (64 and 128 points) the engine carries for
the fast_quad-off rule (src/module/precompute_t.f90:495-891).

Pins: the committed digits equal the nodes/weights the reference itself
builds (oracle/_ref); the 128-point table is Gauss-Legendre on [0,1] mapped
by t = u/(1-u) to 1e-12 in the nodes; the 64-point table (not a textbook
rule: its nodes sit ~4e-5 off that map) still integrates smooth integrands
on [0, inf) to a few 1e-9.
"""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, needs_ref

INC = os.path.join(ROOT, "william_ci_amd", "csrc", "tcg_quad_tables.inc")


def _table(name):
    src = open(INC).read()
    m = re.search(name + r"\[(\d+)\] = \{(.*?)\};", src, re.S)
    vals = np.array([float(x) for x in m.group(2).replace(",", " ").split()])
    assert len(vals) == int(m.group(1))
    return vals


def _gl_mapped(n, s):
    x, w = np.polynomial.legendre.leggauss(n)
    u, wu = 0.5 * (1.0 - x[::-1]), 0.5 * w[::-1]
    return s * u / (1.0 - u), wu * s / (1.0 - u) ** 2


def test_table128_is_mapped_gauss_legendre():
    t, w = _table("kQuad128_t"), _table("kQuad128_w")
    tg, wg = _gl_mapped(128, 1.0)
    assert np.max(np.abs(t / tg - 1.0)) < 2e-12
    assert np.max(np.abs(w / wg - 1.0)) < 5e-11


def test_table64_integrates():
    t, w = _table("kQuad64_t"), _table("kQuad64_w")
    assert np.all(np.diff(t) > 0) and np.all(w > 0)
    tg, _ = _gl_mapped(64, 1.0)
    dev = np.max(np.abs(t / tg - 1.0))
    assert 1e-6 < dev < 1e-4             # the literal table, not the textbook rule
    for f in (lambda x: np.exp(-x), lambda x: 1.0 / (1.0 + x) ** 2,
              lambda x: 2.0 / np.sqrt(np.pi) * np.exp(-x * x)):
        assert abs(np.sum(w * f(t)) - 1.0) < 5e-9


@needs_ref
@pytest.mark.parametrize("n", [64, 128])
def test_tables_equal_reference_nodes(n):
    from oracle import ref
    from william_ci_amd import system as tsys
    r = ref.RefSystem(tsys.workload("h10_ring"), nq=n, fast_quad=False)
    tr, wr, _ = r.nodes()
    assert np.array_equal(_table(f"kQuad{n}_t"), tr)
    assert np.array_equal(_table(f"kQuad{n}_w"), wr)


@needs_ref
def test_oracle_table_rule_matches_reference():
    """The C restatement with the 64-point table as its rule reproduces the
    reference's ERIs on that rule (the rule the reference's self-check picks
    for the H chains)."""
    from oracle import orc, ref
    from william_ci_amd import system as tsys
    s = tsys.workload("h10_ring")
    r = ref.RefSystem(s, nq=64, fast_quad=False)
    t, w, _ = r.nodes()
    Eo, _ = orc.OracleSystem(s, r.coef, nodes=(t, w)).eri_full()
    Er = r.eri_full()
    assert np.abs(Eo - Er).max() <= 1e-13
