"""The oracle's runoptics table loop on the CPU: invariants its structure
fixes (mass0 referencing, the trivial-rh copy, maxrh capping, P
normalisation)."""
import numpy as np
import pytest

from oracle import runoptics as ORO
from tests.runoptics_cases import params, water


@pytest.fixture(scope="module")
def tables():
    return {n: ORO.table(params(n), water()) for n in ("lognorm_simple", "ss_maxrh",
                                                       "bimodal_trivial", "du_trivial")}


def test_dry_mass_reference(tables):
    # at the first rh, area and volume are per dry mass: volume * rhop0 = 1 and
    # area = 3 / (4 rEff rhop0) (area / volume = 3 S2 / (4 S3))
    for name, (ent, nl) in tables.items():
        p = params(name)
        for (b, rhi), v in nl.items():
            if rhi:
                continue
            r0 = p["rhop0"][b] if isinstance(p["rhop0"], list) else p["rhop0"]
            if len(p["psd"]["params"]["fracs"][b]) == 1:
                assert abs(v["volume"] * r0 - 1) < 1e-13
                assert abs(v["area"] * 4 * v["rEff"] * r0 / 3 - 1) < 1e-12


def test_trivial_copies_and_maxrh(tables):
    for name in ("bimodal_trivial", "du_trivial"):
        ent, _ = tables[name]
        for (b, li, rhi), e in ent.items():
            assert e is ent[b, li, 0]
    ent, _ = tables["ss_maxrh"]
    for (b, li, rhi), e in ent.items():
        if rhi == 3:      # rh 0.99 is computed at maxrh = 0.95, the entry before it
            e2 = ent[b, li, 2]
            assert e["growth_factor"] == e2["growth_factor"] and e["qext"] == e2["qext"]
            assert np.array_equal(e["p11"], e2["p11"])


def test_p_normalised(tables):
    from oracle.lut import angle_grid
    th = np.radians(angle_grid())
    for ent, _ in tables.values():
        for e in ent.values():
            assert abs(np.trapezoid(e["p11"] * np.sin(th), th) - 2) < 1e-12
            assert e["pback"][0] == e["p11"][-1]
