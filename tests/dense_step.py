"""A dense FP64 torch restatement of the one-phase DFSPH step, for gradient
checks by torch autograd on tiny scenes.

TEST INFRASTRUCTURE ONLY (like oracle/): imported by tests, never by the
tichi_sph_amd package.

It restates main.py:175-256 with a fixed dt and no FBM (the step the adjoint
differentiates), every pair evaluated at once:
- the sums over neighbours are N x N (fluid, fluid) and N x N_b (fluid,
  boundary) matrices; W, W' and C vanish from r = h on, so a pair outside the
  support adds exactly zero, as the reference's cell walk makes it;
- the surface-tension curvature term has no kernel weight (sph.py:161-166),
  so its pairs are the reference's candidates: the fluid particles in the
  3^d cells around a particle's cell (cell = floor((x - lb) / h),
  sph_util.py:65-72);
- the Jacobi solves run the iteration counts they are given (the recorded
  ones), so the function is the recorded path whose derivative the adjoint
  takes; the psi clamp (sph.py:221-223) and the alpha clamp (sph.py:103-105)
  are where()s, whose derivative is that of the branch taken.
"""
import math

import torch


def _spline(q, sig, h):
    """W(q), W'(q) and the cohesion spline C(q) of sph_util.py:4-37 (W' is
    d W / d r, C carries its 32 / pi / h^3)."""
    one = torch.ones_like(q)
    inner, outer = q <= 0.5, (q > 0.5) & (q < 1)
    u = one - q
    w = torch.where(inner, 6 * (q ** 3 - q ** 2) + 1, torch.where(outer, 2 * u ** 3, 0 * q)) * sig
    wg = torch.where(inner, 6 * (3 * q ** 2 - 2 * q), torch.where(outer, -6 * u ** 2, 0 * q)) * (sig / h)
    c = torch.where(inner, 2 * u ** 3 * q ** 3 - 1 / 64, torch.where(outer, u ** 3 * q ** 3, 0 * q))
    c = c * (32 / math.pi / h ** 3)
    return w, wg, c


class DenseDFSPH:
    """Fluid masses m [n], rest densities rho0 [n] and rest volumes V [n],
    boundary positions xb [nb, d] and masses mb [nb] (the boundary never
    moves, sph.py:265-269), and the constants of the step: h, h^2
    (kernel_h[1..2]), sigma (kernel_sig[d]), gravity g [d], dt, the grid's
    lower bound lb [d] and neighbour range R.  `m` is the state's mass: each
    step ends with SPH_update_mass (sph.py:282-286), m = rho0 V."""

    def __init__(self, m, rho0, V, xb, mb, h, h2, sig, g, dt, lb, R=1):
        self.m, self.rho0, self.V, self.xb, self.mb = m, rho0, V, xb, mb
        self.h, self.h2, self.sig, self.dt, self.R = h, h2, sig, dt, R
        self.g = g
        self.lb = lb

    def _pairs(self, xa, xb):
        """x_ij, r_ij (0 on coincident points, differentiable elsewhere), the
        unit vector x_ij / r_ij (0 where r = 0) and r > 0."""
        xij = xa[:, None, :] - xb[None, :, :]
        r2 = (xij * xij).sum(-1)
        pos = r2 > 0
        rs = torch.sqrt(torch.where(pos, r2, torch.ones_like(r2)))
        r = torch.where(pos, rs, torch.zeros_like(rs))
        unit = torch.where(pos[..., None], xij / rs[..., None], torch.zeros_like(xij))
        return xij, r, unit, pos

    def _solve(self, va, iters, gw_ff, gw_fb, X, Xb, rho, alpha):
        """iters Jacobi iterations of main.py:199-213 / 222-233: IPPE_adv_psi
        (sph.py:196-216), the clamp (219-226) and IPPE_update_vel_adv
        (229-244); boundary vel_adv and psi_adv are 0."""
        dt = self.dt
        for _ in range(iters):
            dv = va[:, None, :] - va[None, :, :]
            psi = (rho - self.rho0) + dt * ((gw_ff * dv).sum(-1) * X[None, :]).sum(1) \
                + dt * ((gw_fb * va[:, None, :]).sum(-1) * Xb[None, :]).sum(1)
            psi = torch.where(psi < 0, torch.zeros_like(psi), psi)
            kap = psi / alpha
            coef = kap[:, None] * X[None, :] + kap[None, :] * X[:, None]
            upd = (coef[..., None] * gw_ff).sum(1) + ((kap[:, None] * Xb[None, :])[..., None] * gw_fb).sum(1)
            va = va - (1 / dt) * upd / self.m[:, None]
        return va

    def step(self, x, v, mu, gamma, n_div, n_inc):
        """One step from fluid positions x and velocities v ([n, d]);
        returns the new (x, v)."""
        h, sig, dt = self.h, self.sig, self.dt
        d = x.shape[1]
        X, Xb = self.m, self.mb  # DFSPH: the alias X is the mass (sph_obj.py:65-73)
        _, r_ff, u_ff, pos_ff = self._pairs(x, x)
        xij_ff = x[:, None, :] - x[None, :, :]
        _, r_fb, u_fb, _ = self._pairs(x, self.xb)
        w_ff, wg_ff, c_ff = _spline(r_ff / h, sig, h)
        w_fb, wg_fb, _ = _spline(r_fb / h, sig, h)
        # SPH_prepare_attr (f,f)(f,b), sph.py:50-64: density, self included
        rho = (w_ff * self.m[None, :]).sum(1) + (w_fb * self.mb[None, :]).sum(1)
        gw_ff = wg_ff[..., None] * u_ff  # grad W_ij, 0 for r = 0
        gw_fb = wg_fb[..., None] * u_fb
        # SPH_prepare_alpha_1 (f,f)(f,b), _2 (f,f), alpha (sph.py:67-105)
        a1 = (X[None, :, None] * gw_ff).sum(1) + (Xb[None, :, None] * gw_fb).sum(1)
        a2 = (wg_ff ** 2 * (X * X / self.m)[None, :]).sum(1)
        alpha = (a1 * a1).sum(1) / self.m + a2
        alpha = torch.where(alpha < 1e-4, torch.full_like(alpha, 1e-4), alpha)
        # divergence solve from vel_adv = vel (sph.py:253-256)
        v = self._solve(v, n_div, gw_ff, gw_fb, X, Xb, rho, alpha)
        # advection (main.py:214-219): gravity, viscosity, surface tension
        Vj = X / rho
        A = v[:, None, :] - v[None, :, :]
        lap = 2 * (2 + d) * Vj[None, :] * wg_ff * (A * xij_ff).sum(-1) / (0.01 * self.h2 + r_ff ** 2)
        acc = self.g[None, :] + (lap[..., None] * u_ff).sum(1) * (mu / self.rho0)[:, None]
        normal = h * (-(Vj[None, :] * wg_ff)[..., None] * u_ff).sum(1)
        cell = torch.floor((x.detach() - self.lb[None, :]) / h)
        cand = ((cell[:, None, :] - cell[None, :, :]).abs() <= self.R).all(-1) & pos_ff
        fac = 2 * self.rho0[:, None] / (rho[:, None] + rho[None, :])
        coh = -gamma * (self.m[None, :] * c_ff)[..., None] * u_ff
        curv = gamma * (normal[:, None, :] - normal[None, :, :])
        st = torch.where(cand[..., None], fac[..., None] * (coh + curv), torch.zeros_like(coh))
        acc = acc + st.sum(1)
        va = v + acc * dt
        # density solve, then SPH_update_mass (one phase: rest density rho0)
        # and SPH_update_pos (sph.py:265-269)
        va = self._solve(va, n_inc, gw_ff, gw_fb, X, Xb, rho, alpha)
        self.m = self.rho0 * self.V
        return x + va * dt, va


def tiny_dambreak(dim, gamma=0.07):
    """(config, scenario) dicts of a dam break small enough for dense pair
    matrices: a block of about 125 fluid particles in the corner of a walled
    tank (270 boundary particles in 2D, 1644 in 3D), one phase, rest density
    1000, mu 0.01, surface tension gamma."""
    if dim == 2:
        gps, fluid_end, box, rt, layers = 0.02, [0.2996, 0.1996], [0.5, 0.4], [0.6, 0.5], 3
    else:
        gps, fluid_end, box, rt, layers = 0.01, [0.0598, 0.0598, 0.0598], [0.12, 0.1, 0.1], [0.16, 0.14, 0.14], 2
    lo = [-3 * gps] * dim
    sb = {
        "sim_env": {"global_part_size": gps, "global_dynamic_viscosity": 0.01, "sim_dim": dim,
                    "sim_space_lb": [-5 * gps] * dim, "sim_space_rt": rt, "gravity": [0, -9.81, 0][:dim],
                    "phase_num": 1, "phase_rest_density": [1000], "phase_color_hex": ["0x6F7DBC"]},
        "fluid": {"max_part_num": 400, "objs": [
            {"type": "cube", "vel": [0] * dim, "particle_relaxing_factor": 1.0, "start_pos": [0.0] * dim,
             "end_pos": fluid_end, "volume_frac": [1], "color": "0x6F7DBC"}]},
        "bound": {"max_part_num": 8000, "objs": [
            {"type": "box", "vel": [0] * dim, "particle_relaxing_factor": 1.0, "start_pos": lo,
             "end_pos": box, "layers": layers, "volume_frac": [1], "color": "0xAAAAAA"}]},
    }
    cb = {"arch": "gpu", "device_memory_GB": 6, "solver_type": "DFSPH", "solver_wc_gamma": 7,
          "solver_divergence_threshold": 0.001, "solver_compression_threshold": 0.001,
          "solver_iter_threshold_min": 2, "solver_iter_threshold_max": 100, "solver_if_cfl": True,
          "solver_cs": 50 if dim == 2 else 100, "solver_cfl_factor": 0.5, "solver_neighb_range": 1,
          "solver_neighb_grid_size_TO_global_part_size": 2, "solver_fbm_diffusion_term": 0.0,
          "solver_fbm_convection_term": 50, "solver_surface_tension_gamma": gamma,
          "gui_res": [1080, 1080], "gui_part_zoomer": 1.0, "gui_fps": 30,
          "gui_canvas_bgcolor": [0.2, 0.2, 0.6], "gui_camera_pos": [6.0, 1.0, 0.0],
          "gui_camera_lookat": [0.0, 0.0, 0.0]}
    return cb, sb
