"""The reference's module-level construction forms (main.py:8-12) on a
bound Config, without a GPU: Ngrid(), Fluid(max_part_num=...) in creation
order (uids 0, 1, sph_obj.py:13-14), NeighborSearch(number_of_uids), Grid(),
config values read as 0-D fields (config.x[None]), and the Simulation that
adopts the objects for scenario ingestion (main.py:62-79)."""
import pytest

from conftest import load_scene


def test_reference_construction_forms_bind_to_one_simulation():
    import tichi_sph_amd as T
    cb, sb = load_scene("dambreak_2d_8k")
    config = T.Config(cb, sb)
    T.use_config(config, device="cpu")
    ngrid = T.Ngrid()
    fluid = T.Fluid(max_part_num=config.fluid_max_part_num[None])
    bound = T.Fluid(max_part_num=config.bound_max_part_num[None])
    assert (fluid.uid, bound.uid) == (0, 1)
    assert fluid.max_part_num == sb["fluid"]["max_part_num"]
    sim = T.current_simulation()
    assert sim.fluid is fluid and sim.bound is bound and fluid.sim is sim and bound.sim is sim
    assert ngrid.sim is sim and T.NeighborSearch(2).sim is sim and T.Grid().shape == (105, 55)
    sim.init_scenario()  # main.py:62-79
    assert (fluid.part_num, bound.part_num) == (128 * 64, 1836)
    # the run loop's 0-D reads (main.py:200, 213, 221, 233)
    assert config.iter_threshold_min[None] == cb["solver_iter_threshold_min"]
    assert config.iter_threshold_max[None] == cb["solver_iter_threshold_max"]
    assert config.divergence_threshold[None] == pytest.approx(cb["solver_divergence_threshold"], rel=1e-7)
    assert config.compression_threshold + 0.0 == config.compression_threshold[None]
    # no new objects once the Simulation exists; no objects without a bound Config
    with pytest.raises(RuntimeError):
        T.Fluid(max_part_num=10)
    from tichi_sph_amd import binding
    binding._CURRENT = None
    with pytest.raises(RuntimeError):
        T.Fluid(max_part_num=10)
    with pytest.raises(RuntimeError):
        T.Ngrid().sim


def test_simulation_refuses_foreign_objects():
    import tichi_sph_amd as T
    cb, sb = load_scene("dambreak_2d_8k")
    c1, c2 = T.Config(cb, sb), T.Config(cb, sb)
    f = T.Fluid(c1, 100, 0, "cpu", 4)
    b = T.Fluid(c1, 100, 1, "cpu", 4)
    T.Simulation(c1, device="cpu", fluid=f, bound=b)
    with pytest.raises(ValueError):
        T.Simulation(c2, device="cpu", fluid=f, bound=b)  # another Config
    with pytest.raises(ValueError):
        T.Simulation(c1, device="cpu", fluid=b, bound=f)  # uids swapped
