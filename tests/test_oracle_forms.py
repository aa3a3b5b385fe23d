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
LJG, IPL12, IPL9, IPL18, IPL6LN, LJLN, GIRI, DZU, FJ, SHOULDER, SAAP, COUL = range(8, 20)


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


# (form, potential factory, SetParams args after (i, j), defining formula v(r),
# rc, smallest r tested)
S, E = 1.1, 0.9
DZU_ARGS = (1.87, 1.94, 1.1, 0.27, 5.82, 1.28)  # a, b, c, d, A, B
FJ_ARGS = (1.0, 1.0, 4.56, 28.88, 1.36, 1.0, 3.57, 2.36, 20.0, 4.0)  # a, eps0, A0, A1, A2, B0, B1, B2, n, Rcut
SAAP_A = (65214.64725, -9.452343340, -19.42488828, -1.958381959, -2.379111084, 1.051490962, -0.1)


def _f32(*x):
    return [float(np.float32(v)) for v in x]


def _dzugutov(r):
    a, b, c, d, A, B = _f32(*DZU_ARGS)
    r = np.asarray(r, dtype=float)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        v2 = np.where(r < b, B * np.exp(d / (r - b)), 0.0)
        v1 = np.where(r < a, A * (r ** -16 - B) * np.exp(c / (r - a)), 0.0)
    return v1 + v2


def _fermi_jagla(r):
    a, e0, A0, A1, A2, B0, B1, B2, n, _ = _f32(*FJ_ARGS)
    x = r / a
    return e0 * (x ** -n + A0 / (1 + np.exp(A1 / A0 * (x - A2))) - B0 / (1 + np.exp(B1 / B0 * (x - B2))))


def _saap(r, sigma=1.0):
    a0, a1, a2, a3, a4, a5, a6 = _f32(*SAAP_A)
    x = r / sigma
    return E * (a0 / x * np.exp(a1 * x + a6 * x * x) + a2 * np.exp(a3 * x) + a4) / (1 + a5 * x ** 6)


def _girifalco(r, alpha=0.5, beta=0.01, r0=0.5):
    x = r / r0
    return (-alpha * (1 / (x * (x - 1) ** 3) + 1 / (x * (x + 1) ** 3) - 2 / x ** 4)
            + beta * (1 / (x * (x - 1) ** 9) + 1 / (x * (x + 1) ** 9) - 2 / x ** 10))


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
        (LJG, lambda: rumd.Pot_LJ_12_6_Gauss(cutoff_method=rumd.NoShift), (S, E, 2.5, 0.15, 1.5, 1.5),
         lambda r: E * ((S / r) ** 12 - 2 * (S / r) ** 6 - 1.5 * np.exp(-(r - 1.5) ** 2 / (2 * 0.15 ** 2))), 2.5 * S),
        (IPL12, lambda: rumd.Pot_IPL_12(cutoff_method=rumd.NoShift), (S, E, 2.5), lambda r: E * (S / r) ** 12, 2.5 * S),
        (IPL9, lambda: rumd.Pot_IPL_9(cutoff_method=rumd.NoShift), (S, E, 2.5), lambda r: E * (S / r) ** 9, 2.5 * S),
        (IPL18, lambda: rumd.Pot_IPL_18(cutoff_method=rumd.NoShift), (S, E, 2.5), lambda r: E * (S / r) ** 18, 2.5 * S),
        (IPL6LN, lambda: rumd.Pot_IPL6_LN(cutoff_method=rumd.NoShift), (0.6, E, 4.0),
         lambda r: E * (0.6 / r) ** 6 / np.log(r / 0.6), 2.4),
        (LJLN, lambda: rumd.Pot_LJ_LN(cutoff_method=rumd.NoShift), (S, 0.5, E, 2.5),
         lambda r: E * ((np.log(S / 0.5) + 1 / 6) * (S / r) ** 12 - (2 * np.log(S / 0.5) + 1 / 6) * (S / r) ** 6)
         / np.log(r / 0.5), 2.5 * S),
        (GIRI, lambda: rumd.Pot_Girifalco(cutoff_method=rumd.NoShift), (0.5, 0.01, 0.5, 5.0), _girifalco, 2.5),
        (DZU, lambda: rumd.Pot_Dzugutov(), DZU_ARGS, _dzugutov, 1.94 - 1e-6),
        (FJ, lambda: rumd.Pot_Fermi_Jagla(cutoff_method=rumd.NoShift), FJ_ARGS, _fermi_jagla, 4.0),
        (SHOULDER, lambda: rumd.Pot_Repulsive_Shoulder(cutoff_method=rumd.NoShift), (S, E, 10.0, 1.35, 2.5),
         lambda r: (S / r) ** 14 + 0.5 * E * (1 - np.tanh(10.0 * (r - 1.35))), 2.5 * S),
        (SAAP, lambda: rumd.Pot_SAAP(cutoff_method=rumd.NoShift), (1.0, E, *SAAP_A, 3.0), _saap, 3.0),
        (COUL, lambda: rumd.Pot_ShiftedForceCoulomb(), (1.0, 2.0),
         lambda r: 1.0 / r + (r - 2.0) / 4.0 - 0.5, 2.0),
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


def test_fixed_exponent_ipl_classes_match_ipl_n(golden_dir):
    """Pot_IPL_12 / 9 / 18 store (1/sigma^2, n eps, eps) where Pot_IPL_n stores
    (1/sigma^2, eps, n); over a two-type fixture they give the same forces,
    energies and virials for every cut-off method."""
    from helpers import load_golden
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    pos4 = configs.pack_pos4(c.positions, c.types)
    rows = [(0, 0, 1.0, 1.0, 2.5), (0, 1, 0.8, 1.5, 2.5), (1, 1, 0.88, 0.5, 2.5)]
    for n, make, form in ((12.0, rumd.Pot_IPL_12, IPL12), (9.0, rumd.Pot_IPL_9, IPL9), (18.0, rumd.Pot_IPL_18, IPL18)):
        for cutoff in (NS, SP, SF):
            fixed = _table(make(CM[cutoff]), rows, 2)
            general = _table(rumd.Pot_IPL_n(n, CM[cutoff]), rows, 2)
            fa, va = oracle.forces_form_brute(pos4, c.box, 2, fixed, form, cutoff)
            fb, vb = oracle.forces_form_brute(pos4, c.box, 2, general, IPLN, cutoff)
            np.testing.assert_allclose(fa, fb, rtol=1e-6, atol=1e-6 * np.abs(fb).max())
            np.testing.assert_allclose(va, vb, rtol=1e-6, atol=1e-6 * np.abs(vb).max())


def test_new_form_invariants():
    # LJ/ln: minimum -eps at Rm
    tab = _table(rumd.Pot_LJ_LN(cutoff_method=rumd.NoShift), [(0, 0, S, 0.5, E, 2.5)])
    u, f = _u_f(LJLN, tab, float(np.float32(S)))
    assert u == pytest.approx(-E, rel=1e-6) and abs(f) < 1e-5
    # LJ-Gauss without the Gaussian well is the 12-6 with its minimum -eps at sigma
    tab = _table(rumd.Pot_LJ_12_6_Gauss(cutoff_method=rumd.NoShift), [(0, 0, S, E, 2.5, 0.15, 0.0, 1.5)])
    u, f = _u_f(LJG, tab, float(np.float32(S)))
    assert u == pytest.approx(-E, rel=1e-6) and abs(f) < 1e-5
    # Dzugutov: the inner term switches off continuously at r = a
    a = float(np.float32(DZU_ARGS[0]))
    tab = _table(rumd.Pot_Dzugutov(), [(0, 0, *DZU_ARGS)])
    below, _ = _u_f(DZU, tab, float(np.nextafter(np.float32(a), np.float32(0))))
    above, _ = _u_f(DZU, tab, float(np.nextafter(np.float32(a), np.float32(3))))
    assert abs(below - above) < 1e-6
    # shifted-force Coulomb: force and energy vanish at Rcut with no cut-off shift
    tab = _table(rumd.Pot_ShiftedForceCoulomb(), [(0, 0, 1.3, 2.2)])
    u, f = _u_f(COUL, tab, float(np.float32(2.2)) * (1 - 1e-7))
    assert abs(u) < 1e-6 and abs(f) < 1e-6


def test_no_shift_only_classes():
    assert rumd.Pot_Dzugutov().GetPotentialForm() == DZU
    assert rumd.Pot_ShiftedForceCoulomb().GetPotentialForm() == COUL
    assert rumd.Pot_Dzugutov().GetID_String() == "potDzu"
    assert rumd.Pot_IPL_9(rumd.ShiftedPotential).GetID_String() == "potIPL9"


def test_saap_force_scale_for_sigma_not_one():
    """The SAAP force factor carries eps/(sigma^2 r) (include/rumd/PairPotential.h:1115),
    so F = -v'(r) / sigma: the true slope for sigma = 1, scaled by 1/sigma
    otherwise.  Both the oracle and the kernels keep that scale."""
    sigma = 1.3
    tab = _table(rumd.Pot_SAAP(cutoff_method=rumd.NoShift), [(0, 0, sigma, E, *SAAP_A, 3.0)])
    for r in (1.2, 1.6, 2.5):
        h = 2e-4
        r0, r1 = float(np.float32(r - h)), float(np.float32(r + h))
        u0, _ = _u_f(SAAP, tab, r0)
        u1, _ = _u_f(SAAP, tab, r1)
        assert u0 == pytest.approx(_saap(r0, float(np.float32(sigma))), rel=2e-6)
        _, f = _u_f(SAAP, tab, float(np.float32(r)))
        assert f == pytest.approx(-(u1 - u0) / (r1 - r0) / sigma, rel=3e-4, abs=1e-7), r
