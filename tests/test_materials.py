"""Material registry (fdtd_amd.materials) against the reference's MaterialMap.

tests/golden/material_map.xml is what the reference's MaterialMap::print_summary
writes for one material of every model the device path covers
(oracle/ref_harness/harness.cpp, scen_material_map).  Reading it must give
those materials; writing them back must give the same text byte for byte; the
resulting material table must step like hand-set materials.
"""
import os

import numpy as np
import pytest

from fdtd_amd.materials import Dispersion, MaterialMap, MaterialPair
from oracle import refcpu
from scenarios import D, E, gauss_table

XML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "material_map.xml")


def test_reads_the_reference_print_summary():
    mm = MaterialMap.read_xml(XML)
    assert sorted(mm) == ["biased fluid", "biased plasma", "conductor", "ferrite", "fluid plasma", "glass",
                          "resonator", "self fluid", "self plasma", "silver", "vacuum"]
    assert mm["vacuum"].spec() == (1.0, 1.0, [])
    assert mm["glass"].spec() == (2.25, 1.0, [])
    assert mm["ferrite"].spec() == (4.5, 3.2, [])
    assert mm["conductor"].spec() == (2.0, 1.0, [(3, 5e5, 0.0, 1.0)])
    assert mm["silver"].spec() == (3.7, 1.0, [(1, 1.37e16, 2.73e13, 1.0)])
    # Lorentz: Delta is read and kept, but (as in the reference's update) not used
    assert mm["resonator"].electric.delta == 0.8
    assert mm["resonator"].spec() == (1.5, 1.0, [(2, 4.1e15, 1.2e14, 1.0)])
    assert mm["biased plasma"].spec() == (1.0, 1.0, [(1, 2e11, 1e9, 1.0)], (0.3, -0.4, 1.25))
    assert mm["self plasma"].spec() == (1.5, 1.0, [(1, 1.5e11, 2e9, 1.0)], None)
    assert mm["fluid plasma"].spec() == {"fluid": 0, "mu": 1.0, "bias": None}
    assert mm["biased fluid"].spec() == {"fluid": 1, "mu": 1.0, "bias": (0.1, 0.2, 0.9)}
    assert mm["self fluid"].spec() == {"fluid": 2, "mu": 1.0, "bias": None}


def test_writes_the_reference_text():
    text = open(XML).read()
    assert MaterialMap.read_xml(text).to_xml() == text


def test_unsupported_models_raise():
    mm = MaterialMap(x=MaterialPair(Dispersion("Anisotropic"), Dispersion("Vacuum")),
                     z=MaterialPair(Dispersion("Constant", K=2.0), Dispersion("MagnetizedDrude", 1.0, 1e15, 1e13)))
    with pytest.raises(NotImplementedError):
        mm["x"].spec()
    with pytest.raises(NotImplementedError, match="magnetic"):
        mm["z"].spec()
    with pytest.raises(ValueError):
        Dispersion("Plasma")


def test_magnetic_poles_from_the_map():
    """A magnetic Drude / Lorentz / Conductive model gives mu_r = K and its pole for
    set_material_magnetic; on the oracle the applied map steps like the poles set by hand."""
    mm = MaterialMap(y=MaterialPair(Dispersion("Constant", K=2.0), Dispersion("Drude", 1.3, 2e14, 1e13)),
                     w=MaterialPair(Dispersion("Vacuum"), Dispersion("Lorentz", 1.5, 3e14, 2e13, 0.4)),
                     c=MaterialPair(Dispersion("Vacuum"), Dispersion("Conductive", 1.2, 50.0)))
    assert mm["y"].spec() == (2.0, 1.3, [])
    assert mm["y"].magnetic_poles() == [(1, 2e14, 1e13, 1.0)]
    assert mm["w"].magnetic_poles() == [(2, 3e14, 2e13, 1.0)]
    assert mm["c"].magnetic_poles() == [(3, 50.0, 0.0, 1.0)]
    nx = ny = 20
    dx = 1e-6
    dt = 0.5 * dx / (refcpu.constants()["c0"] * np.sqrt(2.0))
    matid = np.zeros((1, ny, nx), np.uint8)
    matid[0, 4:9, 4:9], matid[0, 11:16, 4:9], matid[0, 4:9, 11:16] = 1, 2, 3
    runs = []
    for hand in (False, True):
        s = refcpu.Sim("TM", nx, ny, 1, dx, dt)
        if hand:
            s.set_material(1, 2.0, 1.3)
            s.set_material_magnetic(1, [(1, 2e14, 1e13, 1.0)])
            s.set_material(2, 1.0, 1.5)
            s.set_material_magnetic(2, [(2, 3e14, 2e13, 1.0)])
            s.set_material(3, 1.0, 1.2)
            s.set_material_magnetic(3, [(3, 50.0, 0.0, 1.0)])
        else:
            mm.apply(s, {"y": 1, "w": 2, "c": 3})
        s.set_matid(matid)
        s.add_source(0, 2, 10, 10, 0, 8.854e-12 * np.exp(-((np.arange(60) - 15) / 5.0) ** 2))
        s.step(60)
        runs.append([s.get(3, c) for c in (0, 1)] + [s.get(8, 0)])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
    assert np.abs(runs[0][2]).max() > 0


def test_applied_map_steps_like_hand_set_materials():
    mm = MaterialMap.read_xml(XML)
    ids = {"vacuum": 0, "glass": 1, "conductor": 2, "silver": 3, "resonator": 4, "ferrite": 5}
    nx = ny = 24
    dx = 1e-6
    dt = 0.5 * dx / (refcpu.constants()["c0"] * np.sqrt(2.0))
    matid = np.zeros((1, ny, nx), np.uint8)
    for q, (x0, y0) in enumerate([(3, 3), (12, 3), (3, 12), (12, 12), (17, 17)], start=1):
        matid[0, y0:y0 + 5, x0:x0 + 5] = q
    runs = []
    for hand in (False, True):
        s = refcpu.Sim("TM", nx, ny, 1, dx, dt)
        if hand:
            for name, mid in ids.items():
                s.set_material(mid, *mm[name].spec())
        else:
            mm.apply(s, ids)
        s.set_matid(matid)
        s.add_source(D, 2, 11, 11, 0, gauss_table(60, 15, 5))
        s.step(60)
        runs.append(s.get(E, 2))
    np.testing.assert_array_equal(runs[0], runs[1])
    assert np.abs(runs[0]).max() > 0


def test_magnetized_models_apply_in_3d():
    """The map's MagnetizedDrude / SelfMagnetizedDrude entries become magnetized
    Drude materials: a 3-D run equals the one with the materials set by hand,
    and differs from the same plasma without the bias."""
    mm = MaterialMap.read_xml(XML)
    n, dx = 14, 1e-3
    dt = 0.5 * dx / (refcpu.constants()["c0"] * np.sqrt(3.0))
    matid = np.zeros((n, n, n), np.uint8)
    matid[3:7, 3:11, 3:11] = 1
    matid[8:11, 3:11, 3:11] = 2
    runs = []
    for how in ("map", "hand", "unbiased"):
        s = refcpu.Sim("ThreeD", n, n, n, dx, dt)
        if how == "map":
            mm.apply(s, {"vacuum": 0, "biased plasma": 1, "self plasma": 2})
        else:
            s.set_magnetized_drude(1, 1.0, 1.0, 2e11, 1e9, (0.3, -0.4, 1.25) if how == "hand" else (0.0, 0.0, 0.0))
            s.set_magnetized_drude(2, 1.5, 1.0, 1.5e11, 2e9, None)
        s.set_matid(matid)
        s.add_source(D, 2, 7, 7, 7, 3e9 * gauss_table(50, 12, 4))
        s.step(50)
        runs.append(s.get(E, 0))
    np.testing.assert_array_equal(runs[0], runs[1])
    assert np.abs(runs[0] - runs[2]).max() > 1e-3 * np.abs(runs[0]).max()
