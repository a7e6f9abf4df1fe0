"""GPU parity of the per-phase viscosity extension (scenario key
phase_dynamic_viscosity: mu_i = sum_k phi_ik mu_k in the viscous force of
sph.py:114-129; the reference has one global mu, sph_config.py:148).  The
two-phase scene (phase A: 1000 kg/m^3, mu 0.01; phase B: 500 kg/m^3, mu 0.05,
as the two-phase mixture configuration states) is stepped on the fused path
and on the reference surface and compared with the oracle; a run with the
global mu must differ, so the extension is exercised."""
import json

import numpy as np
import pytest

import oracle as O
from conftest import load_scene
from test_gpu_parity import compare_fields

pytestmark = pytest.mark.gpu


def _perturbed(sim, o, seed):
    rng = np.random.default_rng(seed)
    n = sim.fluid.part_num
    vf = o.get(0, "volume_frac").astype(np.float64) + rng.uniform(-0.02, 0.02, (n, 2))
    vf = np.clip(vf, 0, None)
    vf = (vf / vf.sum(1, keepdims=True)).astype(np.float32)
    rho = [np.float32(1000), np.float32(500)]
    rd = (rho[0] * vf[:, 0] + rho[1] * vf[:, 1]).astype(np.float32)
    m = (rd * o.get(0, "rest_volume")).astype(np.float32)
    v = rng.normal(0, 0.05, (n, 3)).astype(np.float32)  # a shear-rich field for the viscous term
    for name, a in (("volume_frac", vf), ("rest_density", rd), ("mass", m), ("vel", v)):
        o.set(0, name, a)
        sim.fluid.from_numpy(name, a)


@pytest.mark.parametrize("path", ["fused", "reference"])
def test_per_phase_viscosity_step_matches_oracle(path):
    from tichi_sph_amd import simulation_from_dicts
    cb, sb = load_scene("two_phase_3d_4k")
    assert sb["sim_env"]["phase_dynamic_viscosity"] == [0.01, 0.05]
    print(f"[case] two_phase_3d_4k per-phase viscosity, {path}", flush=True)
    sims = []
    for env_mu in (True, False):
        s = json.loads(json.dumps(sb))
        if not env_mu:
            del s["sim_env"]["phase_dynamic_viscosity"]
        sim = simulation_from_dicts(cb, s)
        o = O.Oracle(cb, s, 4)
        o.init_scenario()
        _perturbed(sim, o, seed=3)
        if path == "fused":
            st_g = sim.sph_step(stats=True)
            iters = (st_g.div_iter, st_g.incom_iter)
        else:
            di, ii, _, _, _ = sim.sph_step_reference()
            iters = (di, ii)
        st = o.step(O.STEP_CFL)
        assert iters == (st.div_iter, st.incom_iter)
        sims.append((sim, o))
    sim, o = sims[0]
    worst = compare_fields(sim, o, ["pos", "vel", "vel_adv", "sph_compression", "alpha", "acce_adv"],
                           label=f"per-phase mu {path}")
    if path == "reference":
        assert max(worst.values()) == 0  # the reference surface is bit-exact with the oracle
    a_pp = sim.fluid.to_numpy("acce_adv")
    a_gl = sims[1][0].fluid.to_numpy("acce_adv")
    diff = np.abs(a_pp - a_gl).max() / np.abs(a_gl).max()
    print(f"  per-phase vs global mu: largest acceleration difference {diff:.2e} of the largest", flush=True)
    assert diff > 1e-4
