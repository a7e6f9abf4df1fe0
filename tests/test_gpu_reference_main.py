"""The reference's own call sequence against the package: main.py:8-12's
construction (Ngrid(), Fluid(max_part_num=...)), init_scenario
(main.py:62-79), cfl_condition (main.py:348) and main.py:175-256's sph_step
body, with the reference's names and argument order, written out as the
reference writes it.  The sph.py surface is bit-exact with the oracle."""
import numpy as np
import pytest

import oracle as O
from conftest import load_scene
from test_gpu_parity import FIELDS_F, compare_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["dambreak_2d_8k", "dambreak_3d_8k"])
def test_main_py_sph_step_call_sequence_matches_oracle(name):
    import tichi_sph_amd as T
    from tichi_sph_amd.sph import (IPPE_adv_psi, IPPE_adv_psi_init, IPPE_psi_adv_non_negative,
                                   IPPE_update_vel_adv, SPH_advection_gravity_acc,
                                   SPH_advection_surface_tension_acc, SPH_advection_update_vel_adv,
                                   SPH_advection_viscosity_acc, SPH_clean_value, SPH_prepare_alpha,
                                   SPH_prepare_alpha_1, SPH_prepare_alpha_2, SPH_prepare_attr,
                                   SPH_update_energy, SPH_update_mass, SPH_update_pos, SPH_vel_2_vel_adv,
                                   SPH_vel_adv_2_vel, cfl_condition)
    cb, sb = load_scene(name)
    print(f"[case] main.py call sequence, {name}", flush=True)
    config = T.Config(cb, sb)
    T.use_config(config)
    # main.py:8-12
    ngrid = T.Ngrid()
    fluid = T.Fluid(max_part_num=config.fluid_max_part_num[None])
    bound = T.Fluid(max_part_num=config.bound_max_part_num[None])
    globalvar = T.GlobalVariable()
    T.current_simulation().init_scenario()  # main.py:62-79

    def sph_step():  # main.py:175-256
        """ neighbour search """
        ngrid.clear_node()
        ngrid.encode(fluid)
        ngrid.encode(bound)
        ngrid.mem_shift()
        ngrid.fill_node(fluid)
        ngrid.fill_node(bound)
        """ SPH clean value """
        SPH_clean_value(fluid)
        SPH_clean_value(bound)
        """ SPH compute W and W_grad """
        SPH_prepare_attr(ngrid, fluid, fluid)
        SPH_prepare_attr(ngrid, fluid, bound)
        SPH_prepare_attr(ngrid, bound, bound)
        SPH_prepare_attr(ngrid, bound, fluid)
        SPH_prepare_alpha_1(ngrid, fluid, fluid)
        SPH_prepare_alpha_1(ngrid, fluid, bound)
        SPH_prepare_alpha_2(ngrid, fluid, fluid)
        SPH_prepare_alpha_2(ngrid, bound, fluid)
        SPH_prepare_alpha(fluid)
        SPH_prepare_alpha(bound)
        """ IPPE SPH divergence """
        globalvar.div_iter_count = 0
        SPH_vel_2_vel_adv(fluid)
        while globalvar.div_iter_count < config.iter_threshold_min[None] or \
                fluid.compression[None] > config.divergence_threshold[None]:
            IPPE_adv_psi_init(fluid)
            IPPE_adv_psi(ngrid, fluid, fluid)
            IPPE_adv_psi(ngrid, fluid, bound)
            IPPE_psi_adv_non_negative(fluid)
            IPPE_update_vel_adv(ngrid, fluid, fluid)
            IPPE_update_vel_adv(ngrid, fluid, bound)
            globalvar.div_iter_count += 1
            if globalvar.div_iter_count > config.iter_threshold_max[None]:
                break
        SPH_vel_adv_2_vel(fluid)
        """ SPH advection """
        SPH_advection_gravity_acc(fluid)
        SPH_advection_viscosity_acc(ngrid, fluid, fluid)
        SPH_advection_surface_tension_acc(ngrid, fluid, fluid)
        SPH_advection_update_vel_adv(fluid)
        """ IPPE SPH pressure """
        globalvar.incom_iter_count = 0
        while globalvar.incom_iter_count < config.iter_threshold_min[None] or \
                fluid.compression[None] > config.compression_threshold[None]:
            IPPE_adv_psi_init(fluid)
            IPPE_adv_psi(ngrid, fluid, fluid)
            IPPE_adv_psi(ngrid, fluid, bound)
            IPPE_psi_adv_non_negative(fluid)
            IPPE_update_vel_adv(ngrid, fluid, fluid)
            IPPE_update_vel_adv(ngrid, fluid, bound)
            globalvar.incom_iter_count += 1
            if globalvar.incom_iter_count > config.iter_threshold_max[None]:
                break
        """ SPH update """
        SPH_update_mass(fluid)
        SPH_update_pos(fluid)
        SPH_update_energy(fluid)

    o = O.Oracle(cb, sb, 4)
    o.init_scenario()
    for step in range(3):
        cfl_condition(fluid)  # main.py:348
        sph_step()
        st = o.step(O.STEP_CFL)
        assert (globalvar.div_iter_count, globalvar.incom_iter_count) == (st.div_iter, st.incom_iter), step
    worst = compare_fields(T.current_simulation(), o, FIELDS_F, label="main.py sequence")
    print(f"  largest error after 3 steps {max(worst.values()):.2e}", flush=True)
    assert max(worst.values()) == 0
    c = T.current_simulation().control()
    assert c.ke == pytest.approx(o.scalar("ke"), rel=1e-6) and c.pe == pytest.approx(o.scalar("pe"), rel=1e-6)
