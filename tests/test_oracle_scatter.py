"""Pins for the oracle's node scatter (ScatterCell / ScatterCellDiel,
TlmSolver.cpp:59-113) that do not restate its formula:

* the reference self test's own bars (SelfTest.cpp:215-219);
* Johns' conservation laws stated from geometry: charge per polarisation,
  reversal of the loop circulation (Stokes loop in the ij plane), continuity
  of E at the node centre;
* S symmetric orthogonal with entries in {0, +-1/2}, 4 non-zeros per row;
* stub node: weighted unitarity without loss, and Tellegen's energy balance
  with loss (links + stubs lose exactly the power gh * sum_j V_j^2 spent in
  the shunt conductance);
* the checks reject plausible mistakes planted in a Python copy of the node.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle_tlm as ot

# Port numbering is part of the interface (TlmSolver.cpp:20-29):
# p = axis*4 + side*2 + pol, POL[axis][pol] = E axis.
POL = [[1, 2], [0, 2], [0, 1]]
PORTS = [(a, s, pol) for a in range(3) for s in range(2) for pol in range(2)]
EAXIS = np.array([POL[a][pol] for a, s, pol in PORTS])


def port(axis, side, eaxis):
    return axis * 4 + side * 2 + POL[axis].index(eaxis)


def circulation(v, k):
    """Loop integral of the link voltages around axis k (i, j, k cyclic):
    counter-clockwise seen from +k the loop runs +j on the +i face, -i on the
    +j face, -j on the -i face and +i on the -j face; with E = -V/dl the
    circulation in link-voltage units is V(i-,j) - V(i+,j) - V(j-,i) + V(j+,i)."""
    i, j = (k + 1) % 3, (k + 2) % 3
    return 0.5 * (v[port(i, 0, j)] - v[port(i, 1, j)] - v[port(j, 0, i)] + v[port(j, 1, i)])


def node_voltage(v, j):
    """Thevenin: four unit-admittance lines with sources 2*V_in in parallel
    give V_j = (sum of the four incident pulses polarised along j) / 2."""
    return 0.5 * sum(v[p] for p in range(12) if EAXIS[p] == j)


def smatrix(fn):
    S = np.zeros((12, 12))
    for c in range(12):
        e = np.zeros(12)
        e[c] = 1.0
        S[:, c] = fn(e)
    return S


def oracle_air(v):
    return ot.scatter_cell(v, fp64=True)[0]


def check_air_node(fn, rng):
    """All closed-form properties; returns a list of violated ones."""
    bad = []
    S = smatrix(fn)
    if not np.allclose(S, S.T, atol=1e-14):
        bad.append("symmetric")
    if not np.allclose(S.T @ S, np.eye(12), atol=1e-14):
        bad.append("orthogonal")
    if not np.all(np.isin(np.round(S * 2, 12), [-1.0, 0.0, 1.0])):
        bad.append("entries")
    if not np.all((np.abs(S) > 0).sum(axis=1) == 4):
        bad.append("four-per-row")
    for _ in range(20):
        vin = rng.standard_normal(12)
        vout = fn(vin)
        for j in range(3):
            sel = EAXIS == j
            if abs(vin[sel].sum() - vout[sel].sum()) > 1e-12:
                bad.append("charge")
            for a in range(3):
                if a == j:
                    continue
                tot = (vin[port(a, 0, j)] + vout[port(a, 0, j)] +
                       vin[port(a, 1, j)] + vout[port(a, 1, j)])
                if abs(tot - 2 * node_voltage(vin, j)) > 1e-12:
                    bad.append("E-continuity")
        for k in range(3):
            if abs(circulation(vout, k) + circulation(vin, k)) > 1e-12:
                bad.append("circulation")
    return sorted(set(bad))


def test_reference_selftest_bars():
    # SelfTest.cpp:215-219
    assert ot.scatter_unitarity_error() < 1e-5
    assert ot.diel_scatter_unitarity_error(3.0) < 1e-4


def test_air_node_closed_form():
    rng = np.random.default_rng(0)
    assert check_air_node(oracle_air, rng) == []


def test_air_node_float_matches_double_on_exact_inputs():
    # inputs that are small dyadic rationals are scattered exactly in FP32
    rng = np.random.default_rng(1)
    for _ in range(50):
        vin = rng.integers(-64, 64, 12) / 16.0
        a = ot.scatter_cell(vin, fp64=False)[0]
        b = ot.scatter_cell(vin, fp64=True)[0]
        assert np.array_equal(a, b)


def _py_node(vin, sg_flip=None, pol_swap=False, zi_scale=1.0):
    """Python copy of the SCN node with optional planted mistakes."""
    V = [node_voltage(vin, j) for j in range(3)]
    ZI = [circulation(vin, k) * zi_scale for k in range(3)]
    out = np.zeros(12)
    for p, (i, s, pol) in enumerate(PORTS):
        j = POL[i][pol]
        if pol_swap and p == 5:
            j = POL[i][1 - pol]
        k = 3 - i - j if j != i else 0
        sg = (1.0 if (i + 1) % 3 == j else -1.0) * (1.0 if s else -1.0)
        if sg_flip == p:
            sg = -sg
        out[p] = V[j] + sg * ZI[k] - vin[port(i, 1 - s, j)]
    return out


def test_python_copy_passes_and_mutants_fail():
    rng = np.random.default_rng(2)
    assert check_air_node(_py_node, rng) == []
    assert np.allclose(smatrix(_py_node), smatrix(oracle_air))
    # a sign slip on any single port, a swapped polarisation, a wrong 1/2
    for p in range(12):
        assert check_air_node(lambda v, p=p: _py_node(v, sg_flip=p), rng), p
    assert check_air_node(lambda v: _py_node(v, pol_swap=True), rng)
    assert check_air_node(lambda v: _py_node(v, zi_scale=2.0), rng)


def _diel(vin, vs, yh, gh):
    out, vs2 = ot.scatter_cell(vin, vs, yh, gh, diel=True, fp64=True)
    return out, vs2


@pytest.mark.parametrize("epsr", [1.5, 3.0, 10.0])
def test_stub_node_weighted_unitary(epsr):
    yh = 4.0 * (epsr - 1.0)
    S = np.zeros((15, 15))
    for c in range(15):
        vin = np.zeros(12)
        vs = np.zeros(3)
        if c < 12:
            vin[c] = 1.0
        else:
            vs[c - 12] = 1.0
        out, vs2 = _diel(vin, vs, yh, 0.0)
        S[:12, c] = out
        S[12:, c] = vs2
    W = np.diag([1.0] * 12 + [yh] * 3)
    assert np.allclose(S.T @ W @ S, W, atol=1e-12)


def test_stub_node_reduces_to_air():
    rng = np.random.default_rng(3)
    for _ in range(20):
        vin = rng.standard_normal(12)
        out, vs2 = _diel(vin, np.zeros(3), 0.0, 0.0)
        assert np.allclose(out, oracle_air(vin), atol=1e-14)


@pytest.mark.parametrize("epsr,sigma", [(1.0, 0.02), (4.0, 0.05), (9.0, 0.001)])
def test_lossy_node_energy_balance(epsr, sigma):
    """Tellegen per node: sum(Vi^2 - Vr^2) + yh*sum(vs_i^2 - vs_r^2) equals the
    power gh * sum_j V_j^2 dissipated in the shunt conductance, where V_j is
    the node voltage (the value every polarisation-j link sees)."""
    dl = 0.0149896229
    yh = 4.0 * (max(1.0, epsr) - 1.0)
    gh = sigma * dl * 376.730313
    rng = np.random.default_rng(4)
    for _ in range(20):
        vin = rng.standard_normal(12)
        vs = rng.standard_normal(3) if yh > 0 else np.zeros(3)
        out, vs2 = _diel(vin, vs, yh, gh)
        # node voltage from Millman's theorem on the 4 links, stub and conductance
        Vn = [(2 * (vin[EAXIS == j].sum() + yh * vs[j])) / (4 + yh + gh) for j in range(3)]
        lhs = (vin ** 2).sum() - (out ** 2).sum() + yh * ((vs ** 2).sum() - (vs2 ** 2).sum())
        rhs = gh * sum(v * v for v in Vn)
        assert lhs > 0
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs))


def test_port_table_is_a_bijection():
    seen = set()
    for a, s, e in itertools.product(range(3), range(2), range(3)):
        if e == a:
            continue
        seen.add(port(a, s, e))
    assert seen == set(range(12))
