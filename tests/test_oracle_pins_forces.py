"""Pins of the oracle's surface tension, FBM convection, FBM round loop and
CFL velocity branch against closed forms and invariants.

The reference has no tests or fixtures for these functions and Taichi is not
importable here, so each oracle function is pinned by what the mathematics of
sph.py fixes, written out independently below in float64:
  * surface tension (sph.py:132-166): two- and three-particle closed forms of
    the normals, the Akinci cohesion and the curvature term with gamma != 0,
    the 2 rest_psi_i / (psi_i + psi_j) factor and the phase-0 > 0.99 gate on
    i and on j; bulk normals vanish and free-surface normals point out of the
    fluid; with equal masses and rest densities the pair forces conserve
    momentum;
  * SPH_FBM_convect (sph.py:329-363): the drift velocity and the convective
    flux in closed form; no drift in free fall and for single-phase
    particles; the heavier phase drifts along gravity; the flux conserves each
    phase's volume sum_i V_i phi_ik;
  * SPH_FBM_check_tmp and the round loop (sph.py:366-379, main.py:244-250):
    a constructed state in which exactly one particle goes negative, so the
    loop runs two rounds and leaves that particle's fractions unchanged;
  * cfl_condition (sph.py:41-47): dt = min(gps / cs, min_i gps / |v_i| *
    cfl_factor) over |v_i| > 1e-4, evaluated in the precision's arithmetic.
A dropped term, a wrong sign, an index swap (m_i for m_j, psi_j for psi_i,
n_j - n_i) or a gate on one side only fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import load_scene

GRID = (("clear_node", (0,)), ("encode", (0,)), ("encode", (1,)), ("mem_shift", (0,)),
        ("fill_node", (0,)), ("fill_node", (1,)))


def spline(r, h, sig):
    """(W'(r), C(r)) of sph_util.py:4-37 in float64: the cubic spline's
    derivative and the Akinci et al. 2013 cohesion spline."""
    q = r / h
    if q <= 0.5:
        wg = 6 * (3 * q * q - 2 * q)
        c = 2 * (1 - q) ** 3 * q ** 3 - 1 / 64
    elif q < 1:
        wg = -6 * (1 - q) ** 2
        c = (1 - q) ** 3 * q ** 3
    else:
        wg = c = 0.0
    return wg * sig / h, c * 32 / math.pi / h ** 3


def small_oracle(scene, pos, *, cb_over=None, bound_pos=None):
    """An FP64 oracle of the scene's constants holding only the given fluid
    particles and one boundary particle far away (> h) from all of them."""
    cb, sb = load_scene(scene)
    cb = dict(cb, **(cb_over or {}))
    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    o = O.Oracle(cb, sb, 8, fluid_max=n, bound_max=1)
    o.set_part_num(0, n)
    o.set_part_num(1, 1)
    o.set(0, "pos", pos)
    o.set(1, "pos", np.asarray(bound_pos if bound_pos is not None else [pos[0] + 0.1], dtype=np.float64))
    P = o.P
    for obj, m in ((0, n), (1, 1)):
        vf = np.zeros((m, P))
        vf[:, 0] = 1.0
        o.set(obj, "volume_frac", vf)
        o.set(obj, "rest_volume", np.full(m, 1e-6))
        o.set(obj, "rest_density", np.full(m, 1000.0))
        o.set(obj, "mass", np.full(m, 1e-3))
    return o, cb


def run(o, *calls):
    for k, args in calls:
        o.kernel(k, *args)


# ------------------------------------------------------------- surface tension


@pytest.mark.parametrize("q", [0.2, 0.75])
def test_surface_tension_two_particle_closed_form(q):
    """DFSPH, one phase (X = mass, psi = sph_density, rest_psi = rest_density):
    n_i = h sum_j -X_j / psi_j W'(r) x_ij / r, and
    a_i = 2 psi0_i / (psi_i + psi_j) (-gamma m_j C(r) x_ij / r + gamma (n_i - n_j)),
    with asymmetric masses, densities and rest densities so that an index
    swap changes the result.  q = 0.2 has C < 0 (repulsive cohesion),
    q = 0.75 has C > 0 (attractive)."""
    gamma = 0.07
    h0 = 0.02  # the 3D 8k scene: gps 0.01, h = 0.02
    d = np.array([0.6, 0.64, 0.48])  # unit direction
    base = np.array([0.1, 0.1, 0.1])
    o, _ = small_oracle("dambreak_3d_8k", [base, base + q * h0 * d],
                        cb_over={"solver_surface_tension_gamma": gamma})
    run(o, *GRID, ("SPH_clean_value", (0,)))
    m = np.array([1.0e-3, 1.7e-3])
    psi = np.array([950.0, 1230.0])
    psi0 = np.array([1000.0, 880.0])
    o.set(0, "mass", m)
    o.set(0, "sph_density", psi)
    o.set(0, "rest_density", psi0)
    o.kernel("SPH_advection_surface_tension_acc", 0, 0)
    h, sig = o.c.kernel_h[1], o.c.kernel_sig[3]
    x = o.get(0, "pos")
    x01 = x[0] - x[1]
    r = float(np.linalg.norm(x01))
    assert r < h
    e01 = x01 / r
    wg, C = spline(r, h, sig)
    assert (C < 0) == (q < 0.25)
    n0 = h * (-m[1] / psi[1] * wg) * e01
    n1 = h * (-m[0] / psi[0] * wg) * (-e01)
    a0 = 2 * psi0[0] / (psi[0] + psi[1]) * (-gamma * m[1] * C * e01 + gamma * (n0 - n1))
    a1 = 2 * psi0[1] / (psi[1] + psi[0]) * (-gamma * m[0] * C * (-e01) + gamma * (n1 - n0))
    gam32 = float(np.float32(gamma))  # the config stores gamma in f32
    a0 *= gam32 / gamma
    a1 *= gam32 / gamma
    assert o.get(0, "normal")[0] == pytest.approx(n0, rel=1e-12)
    assert o.get(0, "normal")[1] == pytest.approx(n1, rel=1e-12)
    acc = o.get(0, "acce_adv")
    assert acc[0] == pytest.approx(a0, rel=1e-10)
    assert acc[1] == pytest.approx(a1, rel=1e-10)
    # the cohesion part alone attracts at q = 0.75 and repels at q = 0.2
    coh0 = -gam32 * m[1] * C * e01
    assert (coh0 @ (x[1] - x[0]) > 0) == (q > 0.5)


def test_surface_tension_gate_on_i_and_on_j():
    """VFSPH, two phases (X = rest_volume, psi = sph_compression, rest_psi = 1).
    Three mutually close particles with phase-0 fractions 1.0, 0.995 and
    0.985: only the pair (0, 1) passes the gate (both > 0.99).  Particle 2
    fails it as i (its force stays zero although both of its neighbours pass)
    and as j (particles 0 and 1 get no term from it), while the normals of
    loop 1 still sum over all pairs."""
    gamma = 0.07
    base = np.array([0.05, 0.05, 0.05])
    pos = [base, base + [0.006, 0.003, 0.0], base + [0.002, 0.0, 0.007]]
    o, _ = small_oracle("two_phase_3d_4k", pos, cb_over={"solver_surface_tension_gamma": gamma})
    run(o, *GRID, ("SPH_clean_value", (0,)))
    phi0 = np.array([1.0, 0.995, 0.985])
    o.set(0, "volume_frac", np.stack([phi0, 1 - phi0], 1))
    V = np.array([1.0e-6, 1.3e-6, 0.8e-6])
    m = np.array([1.0e-3, 1.25e-3, 0.7e-3])
    psi = np.array([0.97, 1.06, 0.91])
    o.set(0, "rest_volume", V)
    o.set(0, "mass", m)
    o.set(0, "sph_compression", psi)
    o.kernel("SPH_advection_surface_tension_acc", 0, 0)
    h, sig = o.c.kernel_h[1], o.c.kernel_sig[3]
    x = o.get(0, "pos")
    n = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            if i != j:
                xij = x[i] - x[j]
                r = np.linalg.norm(xij)
                assert r < h
                n[i] += -V[j] / psi[j] * spline(r, h, sig)[0] * xij / r
    n *= h
    assert o.get(0, "normal") == pytest.approx(n, rel=1e-12)
    g32 = float(np.float32(gamma))
    a = np.zeros((3, 3))
    for i, j in ((0, 1), (1, 0)):
        xij = x[i] - x[j]
        r = np.linalg.norm(xij)
        C = spline(r, h, sig)[1]
        a[i] = 2 * 1.0 / (psi[i] + psi[j]) * (-g32 * m[j] * C * xij / r + g32 * (n[i] - n[j]))
    acc = o.get(0, "acce_adv")
    assert np.all(acc[2] == 0.0)
    assert acc[:2] == pytest.approx(a[:2], rel=1e-10)


def _lattice_after_density(scene, gamma):
    cb, sb = load_scene(scene)
    o = O.Oracle(dict(cb, solver_surface_tension_gamma=gamma), sb, 8)
    o.init_scenario()
    run(o, *GRID, ("SPH_clean_value", (0,)), ("SPH_prepare_attr", (0, 0)), ("SPH_prepare_attr", (0, 1)))
    return o


def test_surface_tension_normals_bulk_and_free_surface():
    """On the 20^3 lattice block (pid = 400 a + 20 b + c for lattice index
    (a, b, c), y up): a bulk particle's normal vanishes (its neighbourhood is
    point-symmetric), and a particle in the middle of the top face has a
    normal along +y, out of the fluid, with no x or z component."""
    o = _lattice_after_density("dambreak_3d_8k", 0.07)
    o.kernel("SPH_advection_surface_tension_acc", 0, 0)
    nrm = o.get(0, "normal")
    top = nrm[10 * 400 + 19 * 20 + 10]
    bulk = nrm[10 * 400 + 10 * 20 + 10]
    # (the lattice positions are f32-rounded, which breaks the symmetry at
    # the 1e-7 level)
    assert top[1] > 0
    assert abs(top[0]) <= 1e-6 * top[1] and abs(top[2]) <= 1e-6 * top[1]
    assert np.abs(bulk).max() <= 1e-6 * top[1]
    # the magnitude: h times the jump of the colour field across the surface
    # is of order one (a particle on the surface sees about half a kernel)
    assert 0.1 < top[1] < 2.0


def test_surface_tension_conserves_momentum_with_equal_masses():
    """With equal masses and rest densities the pair terms of loop 2 are
    antisymmetric (cohesion along x_ij, curvature n_i - n_j, the factor
    symmetric in psi_i, psi_j), so sum_i m_i a_i = 0 over the fluid block."""
    o = _lattice_after_density("dambreak_3d_8k", 0.07)
    o.kernel("SPH_advection_surface_tension_acc", 0, 0)
    m = o.get(0, "mass")
    assert np.all(m == m[0])
    a = o.get(0, "acce_adv")
    mom = (m[:, None] * a).sum(0)
    assert np.abs(a).max() > 0
    assert np.abs(mom).max() <= 1e-12 * (m[:, None] * np.abs(a)).sum()


# ------------------------------------------------------------- FBM convection


def drift_closed_form(vel, vel_adv, phi, rho0, rho, g, dt):
    """SPH_FBM_convect loop 1 (sph.py:332-346) in float64, per particle."""
    aa = (vel_adv - vel) / dt
    zeta = (phi * (rho[None, :] - rho0[:, None]) / rho[None, :]).sum(1)
    fa = (aa - zeta[:, None] * g[None, :]) / (1 - zeta[:, None])
    u = np.zeros((len(phi), len(rho), len(g)))
    for k in range(len(rho)):
        uk = (phi[:, k] * (rho[k] - rho0))[:, None] * (g[None, :] - fa)
        w = phi[:, k] * rho[k]
        uk = np.where((w > 1e-6)[:, None], uk / np.where(w > 1e-6, w, 1.0)[:, None], uk * 0)
        u[:, k] = (uk + (fa - aa)) * dt
    return aa, zeta, fa, u


def _two_phase_state(seed, n=None):
    cb, sb = load_scene("two_phase_3d_4k")
    o = O.Oracle(cb, sb, 8)
    o.init_scenario()
    n = o.n[0]
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 1, (n, 2))
    phi[: n // 8, 1] = 0.0  # single-phase particles (density weight 0 for phase 1)
    phi /= phi.sum(1, keepdims=True)
    rho = np.array([1000.0, 500.0])
    rho0 = phi @ rho
    o.set(0, "volume_frac", phi)
    o.set(0, "rest_density", rho0)
    o.set(0, "mass", rho0 * o.get(0, "rest_volume"))
    vel = rng.normal(0, 0.1, (n, 3))
    vel_adv = vel + rng.normal(0, 1e-4, (n, 3))
    o.set(0, "vel", vel)
    o.set(0, "vel_adv", vel_adv)
    o.scalar("dt", 0, 1e-4)
    return o, cb, sb, phi, rho, rho0, vel, vel_adv


def test_fbm_drift_velocity_closed_form():
    o, cb, sb, phi, rho, rho0, vel, vel_adv = _two_phase_state(0)
    run(o, *GRID, ("SPH_clean_value", (0,)), ("SPH_FBM_clean_tmp", (0,)), ("SPH_FBM_convect", (0, 0)))
    g = np.array([0.0, float(np.float32(-9.81)), 0.0])
    aa, zeta, fa, u = drift_closed_form(vel, vel_adv, phi, rho0, rho, g, 1e-4)
    assert o.get(0, "acce_adv") == pytest.approx(aa, rel=1e-12, abs=1e-12)
    assert o.get(0, "fbm_zeta") == pytest.approx(zeta, rel=1e-12, abs=1e-15)
    assert o.get(0, "acce") == pytest.approx(fa, rel=1e-12, abs=1e-12)
    got = o.get(0, "drift_vel")
    assert got == pytest.approx(u, rel=1e-9, abs=1e-15)
    assert np.abs(u).max() > 0


def test_fbm_drift_vanishes_in_free_fall_and_for_pure_particles():
    """Free fall (vel_adv - vel = g dt): fbm_acce = g and every phase's drift
    is zero (up to rounding), whatever the fractions; a particle holding one
    phase only has zero drift for any acceleration; at rest (hydrostatic) the
    heavier phase drifts along gravity and the lighter one against it."""
    o, cb, sb, phi, rho, rho0, vel, vel_adv = _two_phase_state(1)
    n = len(phi)
    g = np.array([0.0, float(np.float32(-9.81)), 0.0])
    o.set(0, "vel_adv", vel + g * 1e-4)
    run(o, *GRID, ("SPH_clean_value", (0,)), ("SPH_FBM_convect", (0, 0)))
    u = o.get(0, "drift_vel")
    assert np.abs(u).max() <= 1e-12
    # one phase only: rest density equals the phase density, zeta = 0
    pure = np.zeros((n, 2))
    pure[:, 0] = 1.0
    o.set(0, "volume_frac", pure)
    o.set(0, "rest_density", np.full(n, 1000.0))
    o.set(0, "vel_adv", vel + np.random.default_rng(3).normal(0, 1e-3, (n, 3)))
    o.kernel("SPH_FBM_convect", 0, 0)
    assert np.abs(o.get(0, "drift_vel")).max() <= 1e-15
    # hydrostatic: vel_adv == vel
    o.set(0, "volume_frac", phi)
    o.set(0, "rest_density", rho0)
    o.set(0, "vel_adv", vel)
    o.kernel("SPH_FBM_convect", 0, 0)
    u = o.get(0, "drift_vel")
    mixed = (phi[:, 0] > 0.05) & (phi[:, 1] > 0.05)
    assert mixed.sum() > 100
    assert np.all(u[mixed, 0] @ g > 0)  # phase 0 (1000) is heavier than the mixture
    assert np.all(u[mixed, 1] @ g < 0)


def test_fbm_convection_two_particle_flux_closed_form():
    """tmp_ik -= C dt V_j (phi_ik u_ik + phi_jk u_jk) . (x_ij / r) W'(r), with the
    drift velocities of loop 1, for two particles with different fractions,
    volumes and accelerations (diffusion off)."""
    base = np.array([0.05, 0.05, 0.05])
    pos = [base, base + [0.007, -0.004, 0.002]]
    o, cb = small_oracle("two_phase_3d_4k", pos, cb_over={"solver_fbm_diffusion_term": 0.0})
    phi = np.array([[0.7, 0.3], [0.2, 0.8]])
    rho = np.array([1000.0, 500.0])
    rho0 = phi @ rho
    V = np.array([5.0e-7, 7.0e-7])
    vel = np.array([[0.1, -0.2, 0.05], [-0.3, 0.1, 0.0]])
    vel_adv = vel + np.array([[2e-4, -5e-4, 1e-4], [-1e-4, 3e-4, 0.0]])
    for name, a in (("volume_frac", phi), ("rest_density", rho0), ("rest_volume", V), ("mass", rho0 * V),
                    ("vel", vel), ("vel_adv", vel_adv)):
        o.set(0, name, a)
    dt = 1e-4
    o.scalar("dt", 0, dt)
    run(o, *GRID, ("SPH_clean_value", (0,)), ("SPH_FBM_clean_tmp", (0,)), ("SPH_FBM_convect", (0, 0)))
    g = np.array([0.0, float(np.float32(-9.81)), 0.0])
    _, _, _, u = drift_closed_form(vel, vel_adv, phi, rho0, rho, g, dt)
    h, sig = o.c.kernel_h[1], o.c.kernel_sig[3]
    Cc = float(np.float32(cb["solver_fbm_convection_term"]))
    x = o.get(0, "pos")
    tmp = np.zeros((2, 2))
    for i, j in ((0, 1), (1, 0)):
        xij = x[i] - x[j]
        r = np.linalg.norm(xij)
        wg = spline(r, h, sig)[0]
        for k in range(2):
            f = phi[i, k] * u[i, k] + phi[j, k] * u[j, k]
            tmp[i, k] -= Cc * dt * V[j] * (f @ (xij / r)) * wg
    got = o.get(0, "volume_frac_tmp")
    assert np.abs(tmp).max() > 0
    assert got == pytest.approx(tmp, rel=1e-9)


def test_fbm_convection_conserves_phase_volume():
    """The convective flux is antisymmetric per pair (the bracket is
    symmetric in i, j and x_ij flips sign), so with every particle unflagged
    and diffusion off, sum_i V_i tmp_ik = 0 for each phase."""
    o, cb, sb, phi, rho, rho0, vel, vel_adv = _two_phase_state(2)
    # a separate oracle with diffusion switched off and the same state
    o2 = O.Oracle(dict(cb, solver_fbm_diffusion_term=0.0), sb, 8)
    o2.init_scenario()
    for name in ("volume_frac", "rest_density", "mass", "vel", "vel_adv"):
        o2.set(0, name, o.get(0, name))
    o2.scalar("dt", 0, 1e-4)
    run(o2, *GRID, ("SPH_clean_value", (0,)), ("SPH_FBM_clean_tmp", (0,)), ("SPH_FBM_convect", (0, 0)))
    tmp = o2.get(0, "volume_frac_tmp")
    V = o2.get(0, "rest_volume")
    assert np.abs(tmp).max() > 0
    tot = (V[:, None] * tmp).sum(0)
    assert np.abs(tot).max() <= 1e-12 * (V[:, None] * np.abs(tmp)).sum()


# ----------------------------------------------------- FBM check and round loop


def test_fbm_round_loop_flags_one_particle_and_runs_two_rounds():
    """Three fluid particles in a row, pure diffusion with a large
    coefficient.  Particle 0 (phi = [0.02, 0.98], tiny volume) sits next to
    particle 1 (phi = [1, 0], large volume): its phase-1 fraction would go
    negative in round 1, so SPH_FBM_check_tmp flags it and sets
    general_flag.  Round 2 recomputes the fluxes without particle 0 (as i and
    as j); nothing goes negative, so the loop ends after two rounds, and
    SPH_update_volume_frac leaves particle 0 unchanged.  The expected
    fractions are the closed form of SPH_FBM_diffuse (sph.py:308-326) over the
    unflagged pair (1, 2).  Run through the oracle's whole step
    (main.py:175-256 with the FBM block of 244-250)."""
    D = 1.0
    base = np.array([0.05, 0.05, 0.05])
    pos = [base, base + [0.0096, 0, 0], base + [0.024, 0, 0]]  # h = 0.016: q = 0.6, 0.9, 1.5
    o, cb = small_oracle("two_phase_3d_4k", pos,
                         cb_over={"solver_fbm_diffusion_term": D, "solver_fbm_convection_term": 0.0,
                                  "solver_surface_tension_gamma": 0.0})
    phi = np.array([[0.02, 0.98], [1.0, 0.0], [0.5, 0.5]])
    V = np.array([1e-7, 1e-5, 1e-5])
    rho0 = phi @ np.array([1000.0, 500.0])
    for name, a in (("volume_frac", phi), ("rest_volume", V), ("rest_density", rho0), ("mass", rho0 * V)):
        o.set(0, name, a)
    dt = 1e-4
    o.scalar("dt", 0, dt)
    st = o.step(O.STEP_FBM)
    h, sig = o.c.kernel_h[1], o.c.kernel_sig[3]
    eps = 0.01 * o.c.kernel_h[2]
    Df = float(np.float32(D))
    x = np.asarray(pos)

    def tmp_of(i, partners):
        t = np.zeros(2)
        for j in partners:
            r = np.linalg.norm(x[i] - x[j])
            if r < h:
                t += dt * Df * (phi[i] - phi[j]) * V[j] * r * spline(r, h, sig)[0] / (r * r + eps)
        return t

    # round 1 as the closed form sees it: only particle 0 goes negative
    r1 = [phi[i] + tmp_of(i, [j for j in range(3) if j != i]) for i in range(3)]
    assert r1[0][1] < 0 and (r1[1] >= 0).all() and (r1[2] >= 0).all()
    assert np.linalg.norm(x[0] - x[2]) > h
    assert st.fbm_rounds == 2
    assert list(o.get(0, "flag")) == [1, 0, 0]
    got = o.get(0, "volume_frac")
    assert np.array_equal(got[0], phi[0])
    assert got[1] == pytest.approx(phi[1] + tmp_of(1, [2]), rel=1e-12)
    assert got[2] == pytest.approx(phi[2] + tmp_of(2, [1]), rel=1e-12)
    assert abs(got[1, 1]) > 1e-3  # particle 1 did take up phase 1 from particle 2


# ------------------------------------------------------------------------ CFL


@pytest.mark.parametrize("precision", [4, 8])
def test_cfl_velocity_branch_closed_form(precision):
    """cfl_condition (sph.py:41-47): dt = gps / cs, lowered by every fluid
    particle with |v| > 1e-4 to gps / |v| * cfl_factor, in the precision's
    arithmetic (norm = sqrt(x^2 + y^2 + z^2)).  The velocities are chosen so
    that the velocity branch wins, so the result differs from gps / cs."""
    cb, sb = load_scene("dambreak_3d_8k")
    o = O.Oracle(cb, sb, precision)
    o.init_scenario()
    n = o.n[0]
    T = np.float32 if precision == 4 else np.float64
    rng = np.random.default_rng(5)
    v = rng.normal(0, 20.0, (n, 3)).astype(T)
    v[: n // 2] = 0  # at rest: no contribution
    v[n // 2: n // 2 + 10] = T(5e-5)  # below the 1e-4 threshold
    o.set(0, "vel", v)
    o.kernel("cfl_condition", 0)
    gps = T(o.c.part_size[1])
    cs = T(cb["solver_cs"])
    cfl = T(np.float32(cb["solver_cfl_factor"]))
    dt = gps / cs
    vn = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]).astype(T)
    cand = (gps / vn[vn > T(1e-4)] * cfl).astype(T)
    expect = min(dt, cand.min())
    got = o.scalar("dt")
    assert expect < dt
    assert got == float(expect)


# ------------------------------------------------- per-phase viscosity (extension)


def test_per_phase_viscosity_two_particle_closed_form():
    """The per-phase viscosity extension (scenario key phase_dynamic_viscosity;
    the reference has one global mu, sph_config.py:148): the viscous term of
    sph.py:128-129 with mu replaced by particle i's mu_i = sum_k phi_ik mu_k.
    Two particles with different fractions, so that using mu_j, or the
    global mu, changes the result."""
    base = np.array([0.05, 0.05, 0.05])
    dx = np.array([0.006, -0.003, 0.002])
    o, cb = small_oracle("two_phase_3d_4k", [base, base + dx])
    run(o, *GRID, ("SPH_clean_value", (0,)))
    phi = np.array([[0.8, 0.2], [0.3, 0.7]])
    V = np.array([5e-7, 6e-7])
    psi = np.array([0.98, 1.04])
    rho0 = phi @ np.array([1000.0, 500.0])
    v = np.array([[0.4, -0.1, 0.2], [-0.3, 0.5, 0.0]])
    for name, a in (("volume_frac", phi), ("rest_volume", V), ("sph_compression", psi),
                    ("rest_density", rho0), ("vel", v)):
        o.set(0, name, a)
    o.kernel("SPH_advection_viscosity_acc", 0, 0)
    _, sb = load_scene("two_phase_3d_4k")
    muk = np.array([float(np.float32(m)) for m in sb["sim_env"]["phase_dynamic_viscosity"]])
    assert muk[0] != muk[1]
    h, sig = o.c.kernel_h[1], o.c.kernel_sig[3]
    x = o.get(0, "pos")
    got = o.get(0, "acce_adv")
    for i, j in ((0, 1), (1, 0)):
        xij = x[i] - x[j]
        r = np.linalg.norm(xij)
        wg = spline(r, h, sig)[0]
        Vj = V[j] / psi[j]
        lap = 2 * (2 + 3) * Vj * wg * (xij / r) * ((v[i] - v[j]) @ xij) / (0.01 * o.c.kernel_h[2] + r * r)
        mu_i = phi[i] @ muk
        assert got[i] == pytest.approx(lap * mu_i / rho0[i], rel=1e-10), i


@pytest.mark.parametrize("precision", [4, 8])
def test_per_phase_viscosity_reduces_to_global_for_pure_phases(precision):
    """Particles holding one phase only: mu_i = 1 * mu_k + 0 * mu_other = mu_k
    exactly, so the extension's viscous forces equal, bit for bit, those of a
    global mu = mu_k (the reference's kernel)."""
    cb, sb = load_scene("two_phase_3d_4k")
    import json as _json
    mus = sb["sim_env"]["phase_dynamic_viscosity"]
    sb_glob = _json.loads(_json.dumps(sb))
    del sb_glob["sim_env"]["phase_dynamic_viscosity"]
    sb_glob["sim_env"]["global_dynamic_viscosity"] = mus[1]
    out = []
    for s in (sb, sb_glob):
        o = O.Oracle(cb, s, precision)
        o.init_scenario()
        n = o.n[0]
        pure = np.zeros((n, 2))
        pure[:, 1] = 1.0  # every particle phase B
        o.set(0, "volume_frac", pure)
        o.set(0, "vel", np.random.default_rng(4).normal(0, 0.1, (n, 3)))
        run(o, *GRID, ("SPH_clean_value", (0,)), ("SPH_prepare_attr", (0, 0)), ("SPH_prepare_attr", (0, 1)),
            ("SPH_advection_viscosity_acc", (0, 0)))
        out.append(o.get(0, "acce_adv"))
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[1])
