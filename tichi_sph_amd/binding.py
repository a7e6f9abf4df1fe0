"""The reference's module-level construction, bound to one Config.

The reference builds its objects against a global `config` (sph_config.py)
at import time, in main.py:8-12:

    ngrid = Ngrid()
    fluid = Fluid(max_part_num=config.fluid_max_part_num[None])
    bound = Fluid(max_part_num=config.bound_max_part_num[None])

and passes `ngrid` and the objects to the sph.py kernels.  `use_config`
binds those constructor forms to a Config: Fluid(max_part_num) takes the next
uid (creation order, sph_obj.py:13-14), Ngrid() and NeighborSearch(n) refer to
the Simulation that `simulation()` builds, on first use, from the bound
objects (fluid uid 0, boundary uid 1).
"""

_CURRENT = None
_FACTORY = None  # set by solver.py: builds a Simulation from a Binding


class Binding:
    def __init__(self, config, kw):
        self.config = config
        self.kw = dict(kw)
        self.device = kw.get("device", "cuda")
        self.precision = kw.get("precision", 4)
        self.objs = []
        self.sim = None


def use_config(config, **kw):
    """Bind the reference's constructor forms to `config`; `kw` are the
    Simulation's options (device, precision, fbm, wcsph, cfl, ...)."""
    global _CURRENT
    _CURRENT = Binding(config, kw)
    return _CURRENT


def current():
    if _CURRENT is None:
        raise RuntimeError("no bound Config: call tichi_sph_amd.use_config(config) first (the reference's "
                           "objects are built against its global config, sph_config.py)")
    return _CURRENT


def simulation():
    """The Simulation of the bound objects, built on first use."""
    b = current()
    if b.sim is None:
        if len(b.objs) < 2:
            raise RuntimeError("the bound Config needs its fluid and boundary objects (Fluid(max_part_num) "
                               "twice, main.py:9-10) before the grid or a kernel is used")
        b.sim = _FACTORY(b)
    return b.sim
