"""Pins of the oracle's gas-energy equation, auxiliary per-cell values and ad-hoc ge-ge
Jacobian entry (raja_primordial_network.cpp:1369-1371, 1969-2018, 1859-1867) that do not
go through the oracle's own formulas:

* the energy RHS on states in which only one group of processes is active (one ion and
  the electrons, or molecules without electrons), each expectation built from the tables
  named in the reference's HDF5 file (data/primordial_tables.json), with the optical-depth
  factors and the mass density set to distinct values so that a factor on the wrong term,
  a swapped table, a wrong power of de or H_1, or a wrong redshift factor fails;
* mdensity against the mass density the synthetic generator draws, and the optical-depth
  factors against their closed forms at the tau floor, at tau = 1 and 2, and across the
  H2 threshold density;
* the ad-hoc ge-ge entry against a central difference in T (species fixed) of the two
  processes it keeps (H2 line cooling without the CIE factor and H2 formation heating).
"""
import importlib.util
import json
import os

import numpy as np
import pytest

import orc
from conftest import PKG

spec = importlib.util.spec_from_file_location("mvd_problems", os.path.join(PKG, "problems.py"))
problems = importlib.util.module_from_spec(spec)
spec.loader.exec_module(problems)

NAMES = json.load(open(os.path.join(PKG, "data", "primordial_tables.json")))["order"]
TAB = dict(zip(NAMES, orc.load_tables()))
MH = 1.67e-24
# species slots (raja_primordial_network.cpp:920-929)
H2I, H2II, HI, HII, HM, HeI, HeII, HeIII, DE, GE = range(10)


def rates(T):
    """Named rate values and d/dT at temperature T: log-linear interpolation on the table
    grid [1, 1e5] K with 1024 points (raja_primordial_network.cpp:191-195, 1046-1245; the
    interpolation itself is pinned against the cache by test_rates_at_nodes_equal_table)."""
    dbin = (np.log(1e5) - np.log(1.0)) / 1023
    lT = np.log(T)
    b = min(max(int(lT / dbin), 0), 1022)
    t1, t2 = b * dbin, (b + 1) * dbin
    x = (lT - t1) / (t2 - t1)
    val = {k: a[b] + x * (a[b + 1] - a[b]) for k, a in TAB.items()}
    der = {k: (a[b + 1] - a[b]) / T / (t2 - t1) for k, a in TAB.items()}
    return val, der


def energy_rhs(states, T, z=-1.0, mdensity=3.0, cie=0.7, h2o=0.3):
    """Oracle energy RHS (physical units, before 1/mdensity) at fixed T for each state
    ([n, 10] physical number densities), with scale = 1 and the auxiliary values set."""
    n = len(states)
    c = orc.Chem(n, current_z=z)
    c.array("mdensity")[:] = mdensity
    c.array("inv_mdensity")[:] = 1.0 / mdensity
    c.array("cie_oda")[:] = cie
    c.array("h2_oda")[:] = h2o
    y = np.ascontiguousarray(np.asarray(states, dtype=np.float64).reshape(-1))
    ydot = c.rhs_fixed_T(np.full(n, T), y).reshape(n, 10)
    return ydot[:, GE] * mdensity


def state(**kw):
    s = np.zeros(10)
    idx = {"H2I": H2I, "H2II": H2II, "HI": HI, "HII": HII, "HM": HM, "HeI": HeI,
           "HeII": HeII, "HeIII": HeIII, "de": DE}
    for k, v in kw.items():
        s[idx[k]] = v
    s[GE] = 1.0
    return s


TEMPS = [37.0, 812.5, 9.3e3, 6.1e4]


@pytest.mark.parametrize("T", TEMPS)
def test_energy_rhs_ion_states(T):
    """One ion with electrons at two electron densities: each term's table, its power of
    de and its optical-depth factor (cie on every ionic process) are fixed by the pair."""
    k, _ = rates(T)
    cie = 0.7
    for de in (0.37, 2.9):
        got = energy_rhs([state(HII=1.3, de=de), state(HeI=0.8, de=de), state(HeII=1.7, de=de),
                          state(HeIII=0.6, de=de)], T, cie=cie)
        # H+: recombination and bremsstrahlung, both x de (lines 1369: reHII, brem*(H_2 ...))
        e_hii = -cie * de * 1.3 * (k["reHII_reHII"] + k["brem_brem"])
        # He: collisional ionisation
        e_hei = -cie * de * 0.8 * k["ciHeI_ciHeI"]
        # He+: excitation, ionisation, two recombinations, bremsstrahlung x de, plus the
        # two three-body terms x de^2 (ceHeI, ciHeIS)
        e_heii = -cie * 1.7 * (de * (k["ceHeII_ceHeII"] + k["ciHeII_ciHeII"]
                                     + k["reHeII1_reHeII1"] + k["reHeII2_reHeII2"]
                                     + k["brem_brem"])
                               + de * de * (k["ceHeI_ceHeI"] + k["ciHeIS_ciHeIS"]))
        # He++: recombination and bremsstrahlung with charge^2 = 4
        e_heiii = -cie * de * 0.6 * (k["reHeIII_reHeIII"] + 4.0 * k["brem_brem"])
        want = np.array([e_hii, e_hei, e_heii, e_heiii])
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


@pytest.mark.parametrize("T", TEMPS)
def test_energy_rhs_atomic_hydrogen(T):
    """H with electrons: excitation and ionisation (x de, x cie) plus the H2-formation heating
    0.5 H^3 h2mheat / (ncrn / (H ncrd1) + 1), which carries no optical-depth factor; two H
    densities fix the cube and the critical-density denominator."""
    k, _ = rates(T)
    cie, de = 0.7, 0.45
    for h in (0.9, 4.2):
        got = energy_rhs([state(HI=h, de=de)], T, cie=cie)[0]
        heat = 0.5 * h ** 3 * k["h2formation_h2mheat"] / (
            k["h2formation_ncrn"] / (h * k["h2formation_ncrd1"]) + 1.0)
        want = -cie * de * h * (k["ceHI_ceHI"] + k["ciHI_ciHI"]) + heat
        assert got == pytest.approx(want, rel=1e-13, abs=0)


@pytest.mark.parametrize("T", TEMPS)
def test_energy_rhs_molecular_states(T):
    """H2 without electrons: CIE cooling -2.01588 H2 cieco cie mdensity, and H2 line cooling
    in the Glover & Abel (2008) form H2 cie h2o h2lte / (h2lte / S + 1), S = sum of the
    colliders' gamma tables times their densities; with H present the H2-formation term
    0.5 (-H2 H h2mcool + H^3 h2mheat) / (ncrn / (H2 ncrd2 + H ncrd1) + 1) joins."""
    k, _ = rates(T)
    md, cie, h2o = 3.0, 0.7, 0.3
    h2, h = 2.2, 1.1
    got = energy_rhs([state(H2I=h2), state(H2I=h2, HI=h)], T, mdensity=md, cie=cie, h2o=h2o)
    cie_cool = -2.01588 * h2 * k["cie_cooling_cieco"] * cie * md
    S1 = h2 * k["gloverabel08_gaH2"]
    lte1 = -h2 * cie * h2o * k["gloverabel08_h2lte"] / (k["gloverabel08_h2lte"] / S1 + 1.0)
    S2 = S1 + h * k["gloverabel08_gaHI"]
    lte2 = -h2 * cie * h2o * k["gloverabel08_h2lte"] / (k["gloverabel08_h2lte"] / S2 + 1.0)
    form = 0.5 * (-h2 * h * k["h2formation_h2mcool"] + h ** 3 * k["h2formation_h2mheat"]) / (
        k["h2formation_ncrn"] / (h2 * k["h2formation_ncrd2"] + h * k["h2formation_ncrd1"]) + 1.0)
    np.testing.assert_allclose(got, [cie_cool + lte1, cie_cool + lte2 + form],
                               rtol=1e-13, atol=0)


def test_energy_rhs_line_cooling_colliders():
    """Each collider enters S with its own gamma table: adding one collider (H+, He, e-)
    to an H2 gas changes S by exactly that collider's density times its table."""
    T = 2.4e3
    k, _ = rates(T)
    cie, h2o, md = 0.7, 0.3, 3.0
    h2 = 1.5
    base = -2.01588 * h2 * k["cie_cooling_cieco"] * cie * md
    for name, kw, ga, extra in (
            ("H+", dict(HII=0.8), "gloverabel08_gaHp", lambda de: 0.0),
            ("He", dict(HeI=0.8), "gloverabel08_gaHe", lambda de: 0.0)):
        got = energy_rhs([state(H2I=h2, **kw)], T, mdensity=md, cie=cie, h2o=h2o)[0]
        S = h2 * k["gloverabel08_gaH2"] + 0.8 * k[ga]
        want = base - h2 * cie * h2o * k["gloverabel08_h2lte"] / (k["gloverabel08_h2lte"] / S + 1.0)
        assert got == pytest.approx(want, rel=1e-13, abs=0), name
    # electrons also add their own cooling terms; use H2 + e- only (no ion): S gains de*gael
    de = 0.25
    got = energy_rhs([state(H2I=h2, de=de)], T, mdensity=md, cie=cie, h2o=h2o)[0]
    S = h2 * k["gloverabel08_gaH2"] + de * k["gloverabel08_gael"]
    want = base - h2 * cie * h2o * k["gloverabel08_h2lte"] / (k["gloverabel08_h2lte"] / S + 1.0)
    assert got == pytest.approx(want, rel=1e-13, abs=0)


def test_energy_rhs_compton_term():
    """Compton cooling off the CMB, -cie compton de (z+1)^4 (T - 2.73 (z+1)): zero at the
    demo's z = -1; at z = 2 and z = 0.5 and two temperatures it fixes the power and the
    CMB temperature factor."""
    cie, de, hii = 0.7, 0.5, 0.5
    for T in (150.0, 3.3e3):
        k, _ = rates(T)
        ions = -cie * de * hii * (k["reHII_reHII"] + k["brem_brem"])
        assert energy_rhs([state(HII=hii, de=de)], T, z=-1.0, cie=cie)[0] == pytest.approx(
            ions, rel=1e-13, abs=0)
        for z in (2.0, 0.5):
            got = energy_rhs([state(HII=hii, de=de)], T, z=z, cie=cie)[0]
            comp = -cie * k["compton_comp_"] * de * (z + 1.0) ** 4 * (T - 2.73 * (z + 1.0))
            assert got == pytest.approx(ions + comp, rel=1e-12, abs=0), (T, z)


def test_energy_rhs_is_divided_by_mdensity_and_scaled():
    """The stored ge rate is the physical rate x inv_scale[ge] x inv_mdensity (lines
    1370-1371): with scale[ge] = 4 and mdensity = 3 a state's ydot[ge] is 1/12 of the
    physical rate computed above."""
    T = 812.5
    c = orc.Chem(1)
    c.array("mdensity")[:] = 3.0
    c.array("inv_mdensity")[:] = 1.0 / 3.0
    c.array("cie_oda")[:] = 0.7
    c.array("h2_oda")[:] = 0.3
    c.array("scale")[GE] = 4.0
    c.array("inv_scale")[GE] = 0.25
    s = state(HeI=0.8, de=0.37)
    s[GE] = 0.25                      # scaled ge; the energy equation does not read ge
    ydot = c.rhs_fixed_T(np.array([T]), s.copy())
    k, _ = rates(T)
    phys = -0.7 * 0.37 * 0.8 * k["ciHeI_ciHeI"]
    assert ydot[GE] == pytest.approx(phys / 12.0, rel=1e-14, abs=0)


def test_mdensity_is_the_generator_density():
    """mdensity = mH sum_s A_s n_s from the initial (scale) state equals the mass density the
    synthetic generator drew (rho = n_tot mH, split by mass fractions), also after the H-
    cap moved mass to H."""
    n = 512
    y, _, rho = problems.chem_cells(n, seed=17, with_rho=True)
    c = orc.Chem(n)
    yy = y.copy()
    c.prepare(yy)
    np.testing.assert_allclose(c.array("mdensity"), rho, rtol=1e-14, atol=0)
    np.testing.assert_allclose(c.array("inv_mdensity") * rho, 1.0, rtol=1e-14, atol=0)


def _extra_for_density(rho):
    """Auxiliary values of cells of pure atomic hydrogen of mass density rho."""
    rho = np.asarray(rho, dtype=np.float64)
    c = orc.Chem(len(rho))
    sc = c.array("scale")
    sc[:] = 0.0
    sc.reshape(-1, 10)[:, HI] = rho / (1.00794 * MH)
    c.extra()
    return c


def test_optical_depth_factors_closed_forms():
    """cie_oda = min(1, (1 - e^-tau)/tau), tau = max((rho / 3.3e-8)^2.8, 1e-5), and
    h2_oda = min(1, (rho / 1.34e-14)^-0.45) (Ripamonti & Abel 2004 and the H2 line opacity
    fit the network uses), at densities that put tau on its floor, at 1 and at 2, and
    rho below / at / above the H2 threshold."""
    rho_tau1 = 3.3e-8
    rho_tau2 = 3.3e-8 * 2.0 ** (1.0 / 2.8)
    rho = np.array([1e-22, rho_tau1, rho_tau2, 1.34e-15, 1.34e-13, 1.34e-12])
    c = _extra_for_density(rho)
    np.testing.assert_allclose(c.array("mdensity"), rho, rtol=1e-15)
    cie = c.array("cie_oda")
    floor = -np.expm1(-1e-5) / 1e-5
    assert cie[0] == pytest.approx(floor, rel=1e-12)
    assert cie[3] == pytest.approx(floor, rel=1e-12)      # (1e-7)^2.8 is below the floor
    assert cie[1] == pytest.approx(1.0 - np.exp(-1.0), rel=1e-12)
    assert cie[2] == pytest.approx((1.0 - np.exp(-2.0)) / 2.0, rel=1e-12)
    h2o = c.array("h2_oda")
    assert h2o[0] == 1.0 and h2o[1] == pytest.approx(
        (rho_tau1 / 1.34e-14) ** -0.45, rel=1e-12)
    assert h2o[3] == 1.0                                   # below the threshold: capped
    assert h2o[4] == pytest.approx(10.0 ** -0.45, rel=1e-12)
    assert h2o[5] == pytest.approx(100.0 ** -0.45, rel=1e-12)


def test_mdensity_species_weights():
    """Each species counts with its own mass: 2 for H2 and H2+, 1.00794 for H, H+, H-,
    4.002602 for He, He+, He++, nothing for the electrons and the energy."""
    c = orc.Chem(10)
    sc = c.array("scale").reshape(10, 10)
    sc[:] = 0.0
    for s in range(10):
        sc[s, s] = 1.0
    c.extra()
    want = np.array([2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602, 4.002602, 4.002602,
                     0.0, 0.0]) * MH
    np.testing.assert_allclose(c.array("mdensity"), want, rtol=1e-15, atol=0)


def _two_terms(y, T):
    """The processes the ad-hoc ge-ge entry keeps, physical units, species fixed: H2 line
    cooling without the CIE factor and H2 formation heating (raja_primordial_network.cpp:
    1862-1864 comment), with the h2 optical depth factor 1 (applied by the caller)."""
    k, _ = rates(T)
    S = (y[H2I] * k["gloverabel08_gaH2"] + y[HI] * k["gloverabel08_gaHI"]
         + y[HII] * k["gloverabel08_gaHp"] + y[HeI] * k["gloverabel08_gaHe"]
         + y[DE] * k["gloverabel08_gael"])
    lte = -y[H2I] * k["gloverabel08_h2lte"] / (k["gloverabel08_h2lte"] / S + 1.0)
    D = y[H2I] * k["h2formation_ncrd2"] + y[HI] * k["h2formation_ncrd1"]
    form = 0.5 * (-y[H2I] * y[HI] * k["h2formation_h2mcool"]
                  + y[HI] ** 3 * k["h2formation_h2mheat"]) / (k["h2formation_ncrn"] / D + 1.0)
    return lte, form


def _dense_gas_cells():
    """Dense molecular gas (n_H 1e8-1e10 cm^-3) in which H2 line cooling, H2 formation
    cooling (-H2 H h2mcool) and formation heating (H^3 h2mheat) are of similar size, at
    temperatures between 300 K and 2e4 K."""
    rows = []
    for nh in (1e8, 1e9, 1e10):
        for f2 in (1e-3, 0.05, 0.5):
            for T in (320.0, 2.9e3, 1.7e4):
                s = np.zeros(10)
                s[H2I] = f2 * nh
                s[HI] = nh
                s[HII] = 1e-4 * nh
                s[HeI] = 0.08 * nh
                s[HeII] = 1e-6 * nh
                s[DE] = (s[HII] + s[HeII]) * MH / problems.M_AMU
                s[H2II] = s[HM] = s[HeIII] = 1e-12 * nh
                rho = MH * (2.0 * s[H2I] + 1.00794 * s[HI] + 4.002602 * s[HeI])
                s[GE] = problems.KBOLTZ * T * s[:9].sum() * 1.5 / rho
                rows.append(s)
    return np.ascontiguousarray(np.array(rows).reshape(-1))


@pytest.mark.parametrize("cells", ["generator", "dense"])
def test_adhoc_ge_ge_entry_is_dT_of_line_cooling_and_formation(cells):
    """J[ge,ge] = (d/dT)[h2o * lte + form] * Tge * inv_mdensity at the cell's temperature,
    species fixed, with the rates' T-slopes of the log-linear interpolation (central
    difference, h = 1e-6 T, inside one table bin): on the synthetic cells, where line
    cooling dominates, and on dense gas, where the formation terms are of its size."""
    if cells == "generator":
        n = 64
        y, _ = problems.chem_cells(n, seed=29, temp=(30.0, 8e3), alpha=(4.0,) * 8)
    else:
        y = _dense_gas_cells()
        n = len(y) // 10
    c = orc.Chem(n)
    yy = y.copy()
    c.prepare(yy)
    c.rhs(yy)
    J = c.jac(yy)
    Ts = c.array("Ts").copy()
    Tge = c.array("dTs_ge").copy()
    imd = c.array("inv_mdensity").copy()
    h2o = c.array("h2_oda").copy()
    phys = (yy * c.array("scale")).reshape(n, 10)
    dbin = (np.log(1e5) - np.log(1.0)) / 1023
    checked = 0
    for i in range(n):
        T = Ts[i]
        hT = 1e-6 * T
        if int(np.log(T + hT) / dbin) != int(np.log(T - hT) / dbin):
            continue                     # the difference would cross a table node
        lp, fp = _two_terms(phys[i], T + hT)
        lm, fm = _two_terms(phys[i], T - hT)
        d = (h2o[i] * (lp - lm) + (fp - fm)) / (2 * hT)
        want = d * Tge[i] * imd[i]
        scale = abs(h2o[i] * (lp - lm) / (2 * hT)) + abs((fp - fm) / (2 * hT))
        assert abs(J[i, GE, GE] - want) <= 1e-7 * scale * abs(Tge[i] * imd[i]), i
        checked += 1
    assert checked > n // 2
    if cells == "dense":
        # the formation terms matter here: dropping them would fail the comparison
        lp, fp = _two_terms(phys[0], Ts[0] * (1 + 1e-6))
        lm, fm = _two_terms(phys[0], Ts[0] * (1 - 1e-6))
        assert abs(fp - fm) > 1e-3 * abs(h2o[0] * (lp - lm))
