"""Closed-form checks of the CPU oracle (furnace_test.cpp): white furnace,
bounded energy, Kulla-Conty energy conservation and Beer-Lambert slabs.
These pin the oracle's integrator and materials independently of the
golden vectors: a dropped MIS term, a wrong pdf or a missing
compensation lobe moves the furnace mean away from 1."""
import pytest

from furnace_cases import BOUNDED, ENERGY, FURNACE, SLABS, bounded_check, object_energy, run_furnace, slab_worst
from oracle_lib import OracleScene


def oracle_render(b, cam):
    desc, keep = b.build()
    o = OracleScene(desc)
    img = o.render(cam, seed=1, threads=8)["radiance"]
    del keep
    return img


@pytest.mark.parametrize("name,build,check", FURNACE, ids=[c[0] for c in FURNACE])
def test_furnace(name, build, check):
    ok, msg = check(run_furnace(oracle_render, build))
    assert ok, f"{name}: {msg}"


@pytest.mark.parametrize("name,build", BOUNDED, ids=[c[0] for c in BOUNDED])
def test_bounded(name, build):
    ok, msg = bounded_check(run_furnace(oracle_render, build))
    assert ok, f"{name}: {msg}"


@pytest.mark.parametrize("name,make", ENERGY, ids=[c[0] for c in ENERGY])
def test_energy_conserving(name, make):
    e = object_energy(oracle_render, make)
    assert abs(e - 1) < 0.02, f"{name}: sphere pixels {e}"


@pytest.mark.parametrize("name,build,expect", SLABS, ids=[c[0] for c in SLABS])
def test_slab(name, build, expect):
    assert slab_worst(oracle_render, build, expect) < 1e-5


def test_planted_mistake_is_caught():
    """A furnace check must fail for an energy-losing material: a 0.9-albedo
    sphere pulls the mean below 1 - 1e-6."""
    from furnace_cases import furnace_check, sphere_case
    ok, _ = furnace_check(1e-6)(run_furnace(oracle_render, sphere_case(lambda b: b.lambertian((0.9, 0.9, 0.9)))))
    assert not ok
