"""Pins the oracle's pair forms outside the 12-6 family
(oracle/rumd_oracle.c orc_forces_form_brute) and the host's parameter rows
for them: each form's energy against its defining formula in the physical
parameters, -du/dr = F by finite differences, the special cases that reduce
to the pinned 12-6 path, and the invariants the forms are built with.  CPU
only."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import ka_table
from rumd_amd import configs
import rumd_amd.rumd as rumd

NS, SP, SF = 0, 1, 2
CM = {NS: rumd.NoShift, SP: rumd.ShiftedPotential, SF: rumd.ShiftedForce}
SMOOTH, GLJ, GAUSS, BUCK, IPLN, YUK, EXP = 1, 2, 3, 4, 5, 6, 7


def _table(pot, rows, nt=1):
    pot.SetVerbose(False)
    for args in rows:
        pot.SetParams(*args)
    return np.asarray(pot.GetParamsTable(nt), dtype=np.float32)


def _pair(r, L=30.0):
    pos4 = configs.pack_pos4(np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]], np.float32), np.zeros(2, np.int32))
    return pos4, np.array([L, L, L])


def _u_f(form, tab, r, cutoff=NS):
    pos4, L = _pair(r)
    f4, _ = oracle.forces_form_brute(pos4, L, 1, tab, form, cutoff)
    return f4[0, 3] + f4[1, 3], f4[1, 0]  # pair energy; force on particle 1 along +x


# (form, potential factory, SetParams args after (i, j), defining formula v(r))
S, E = 1.1, 0.9


def _cases():
    xc = 2.5
    c2 = 48 * xc ** -14 - 15 * xc ** -8
    c4 = 6 * xc ** -10 - 21 * xc ** -16
    m, n = 9.0, 5.0
    rm, alpha = 1.2, 14.0
    return [
        (SMOOTH, lambda: rumd.Pot_LJ_12_6_smooth_2_4(cutoff_method=rumd.NoShift), (S, E, xc),
         lambda r: 4 * E * ((S / r) ** 12 - (S / r) ** 6 + c2 * (r / S) ** 2 + c4 * (r / S) ** 4), xc * S),
        (GLJ, lambda: rumd.Pot_gLJ_m_n(set_m=m, set_n=n, cutoff_method=rumd.NoShift), (S, E, 2.5),
         lambda r: E / (m - n) * (n * (S / r) ** m - m * (S / r) ** n), 2.5 * S),
        (GAUSS, lambda: rumd.Pot_Gauss(cutoff_method=rumd.NoShift), (S, E, 2.5),
         lambda r: E * np.exp(-(r / S) ** 2), 2.5 * S),
        (BUCK, lambda: rumd.Pot_Buckingham(cutoff_method=rumd.NoShift), (rm, alpha, E, 3.0),
         lambda r: E * (6 / (alpha - 6) * np.exp(alpha * (1 - r / rm)) - alpha / (alpha - 6) * (rm / r) ** 6), 3.0),
        (IPLN, lambda: rumd.Pot_IPL_n(n=18.0, cutoff_method=rumd.NoShift), (S, E, 2.5),
         lambda r: E * (S / r) ** 18, 2.5 * S),
        (YUK, lambda: rumd.Pot_Yukawa(cutoff_method=rumd.NoShift), (S, E, 3.0),
         lambda r: E * S * np.exp(-r / S) / r, 3.0 * S),
        (EXP, lambda: rumd.Pot_Exp(cutoff_method=rumd.NoShift), (S, E, 3.0),
         lambda r: E * np.exp(-r / S), 3.0 * S),
    ]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: str(c[0]))
def test_form_energy_is_the_defining_formula(case):
    form, make, args, v, rc = case
    tab = _table(make(), [(0, 0, *args)])
    assert tab[0, 0] == pytest.approx(rc, rel=1e-6)
    for r in np.linspace(0.9, 0.97 * rc, 9):
        r32 = float(np.float32(r))
        u, _ = _u_f(form, tab, r32)
        assert u == pytest.approx(v(r32), rel=2e-6, abs=1e-9), (form, r32)
    # the host's shift constants: u(rc) and s(rc) rc with s = -u'(r)/r
    h = 1e-6 * rc
    dvdr = (v(rc + h) - v(rc - h)) / (2 * h)
    assert tab[0, 11] == pytest.approx(v(rc), rel=1e-5, abs=1e-9)
    assert tab[0, 10] == pytest.approx(-dvdr, rel=1e-4, abs=1e-7)


@pytest.mark.parametrize("cutoff", [NS, SP, SF])
@pytest.mark.parametrize("case", _cases(), ids=lambda c: str(c[0]))
def test_form_force_is_minus_the_energy_slope(case, cutoff):
    form, make, args, v, rc = case
    tab = _table(make(), [(0, 0, *args)])
    for r in (0.95, 1.3, 0.8 * rc):
        h = 2e-4
        r0, r1 = float(np.float32(r - h)), float(np.float32(r + h))
        u0, _ = _u_f(form, tab, r0, cutoff)
        u1, _ = _u_f(form, tab, r1, cutoff)
        _, f = _u_f(form, tab, float(np.float32(r)), cutoff)
        assert f == pytest.approx(-(u1 - u0) / (r1 - r0), rel=3e-4, abs=1e-7), (form, r, cutoff)
    if cutoff == SF:  # shifted force: force and energy vanish at rc
        u, f = _u_f(form, tab, float(np.float32(rc)) * (1 - 1e-6), SF)
        assert abs(f) < 1e-4 and abs(u) < 1e-6


def test_special_cases_reduce_to_the_12_6_path(golden_dir):
    """gLJ with m = 12, n = 6 is Pot_gLJ_m12_n6 (v = eps ((s/r)^12 - 2 (s/r)^6));
    IPL_n with n = 12 is IPL 12-6 with A12 = eps sigma^12, A6 = 0.  Over a
    two-type fixture, every cut-off method."""
    from helpers import load_golden
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    pos4 = configs.pack_pos4(c.positions, c.types)
    rows = [(0, 0, 1.0, 1.0, 2.5), (0, 1, 0.8, 1.5, 2.5), (1, 1, 0.88, 0.5, 2.5)]
    for cutoff in (NS, SP, SF):
        glj = _table(rumd.Pot_gLJ_m_n(12.0, 6.0, CM[cutoff]), rows, 2)
        lja = _table(rumd.Pot_gLJ_m12_n6(CM[cutoff]), rows, 2)
        fa, va = oracle.forces_form_brute(pos4, c.box, 2, glj, GLJ, cutoff)
        fb, vb = oracle.forces_brute(pos4, c.box, 2, lja, cutoff)
        np.testing.assert_allclose(fa, fb, rtol=2e-6, atol=2e-5)
        np.testing.assert_allclose(va, vb, rtol=2e-6, atol=2e-5)
        ipl = _table(rumd.Pot_IPL_n(12.0, CM[cutoff]), rows, 2)
        ipl126 = rumd.Pot_IPL_12_6(CM[cutoff])
        ipl126_rows = [(i, j, e * s ** 12, 0.0, rc * s) for (i, j, s, e, rc) in rows]
        t126 = _table(ipl126, ipl126_rows, 2)
        fa, _ = oracle.forces_form_brute(pos4, c.box, 2, ipl, IPLN, cutoff)
        fb, _ = oracle.forces_brute(pos4, c.box, 2, t126, cutoff)
        np.testing.assert_allclose(fa, fb, rtol=2e-6, atol=2e-5)


def test_form_invariants():
    # Buckingham: minimum -eps at rm (zero force there)
    rm, alpha = 1.2, 14.0
    tab = _table(rumd.Pot_Buckingham(cutoff_method=rumd.NoShift), [(0, 0, rm, alpha, E, 3.0)])
    u, f = _u_f(BUCK, tab, float(np.float32(rm)))
    assert u == pytest.approx(-E, rel=1e-6) and abs(f) < 1e-5
    # smooth LJ: force zero at Rcut without any shift
    tab = _table(rumd.Pot_LJ_12_6_smooth_2_4(cutoff_method=rumd.NoShift), [(0, 0, S, E, 2.5)])
    _, f = _u_f(SMOOTH, tab, float(np.float32(2.5 * S)) * (1 - 1e-7))
    assert abs(f) < 1e-6
    # generalised LJ: minimum -eps at sigma
    tab = _table(rumd.Pot_gLJ_m_n(9.0, 5.0, rumd.NoShift), [(0, 0, S, E, 2.5)])
    u, f = _u_f(GLJ, tab, float(np.float32(S)))
    assert u == pytest.approx(-E, rel=1e-6) and abs(f) < 1e-5


def test_form_potentials_refuse_the_fp64_path():
    pot = rumd.Pot_Gauss(cutoff_method=rumd.ShiftedPotential)
    with pytest.raises(RuntimeError, match="12-6 potentials only"):
        pot.SetFP64Forces(True)
    assert rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce).GetPotentialForm() == 0
    assert ka_table().shape == (4, 12)
