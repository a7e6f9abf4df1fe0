"""The reference's NeighborSearch surface (NeighborSearch.py:5-132).

`NeighborSearch(number_of_uids)` (the reference's form, on the Config bound
by tichi_sph_amd.use_config) or `NeighborSearch(sim, number_of_uids)` builds,
for every particle of each
object, the candidate list of the reference's fill_neighbs
(NeighborSearch.py:84-101): every particle of the neighbouring cells, self
included and unfiltered by distance, grouped by uid, in template order then
cell order.  `parts[uid].neighbs` is stored as a flat array with row starts
(the reference sizes rows by an undefined max_neighb_per_part constant);
`neighbor_first/last(obj, nobj, i)` give the range of uid `nobj.uid` inside
particle i's row (NeighborSearch.py:115-121).  Entries are storage slots of
the neighbour object.
"""
import ctypes

import torch

from . import _lib


class PartNeighbs:
    def __init__(self):
        self.neighbs = None        # int32 [total]
        self.row_start = None      # int32 [n + 1]
        self.neighbs_shift = None  # int32 [n, number_of_uids]
        self.neighbs_count = None  # int32 [n, number_of_uids]


class NeighborSearch:
    def __init__(self, sim=None, number_of_uids=2):
        if isinstance(sim, int):  # NeighborSearch(number_of_uids), NeighborSearch.py:13
            sim, number_of_uids = None, sim
        self._sim = sim
        self.number_of_uids = number_of_uids
        self.parts = [PartNeighbs() for _ in range(number_of_uids)]

    @property
    def sim(self):
        if self._sim is None:
            from . import binding
            self._sim = binding.simulation()
        return self._sim

    def establish_neighbs(self, *objs):
        sim = self.sim
        g = sim.ngrid
        g.build(*objs)
        U = self.number_of_uids
        P = ctypes.byref(sim.params)
        for o in objs:
            n = o.part_num
            part = self.parts[o.uid]
            dev = sim.device
            shift = torch.zeros(max(n, 1), U, dtype=torch.int32, device=dev)
            count = torch.zeros(max(n, 1), U, dtype=torch.int32, device=dev)
            npc, nps, pin, uin = g.ptrs()
            rc = sim.lib.tsph_neighbor_lists(P, ctypes.byref(o.struct()), U, npc, nps, pin, uin, None, None,
                                             shift.data_ptr(), count.data_ptr(), sim.stream_ptr())
            _lib.check(rc, "tsph_neighbor_lists(count)")
            tot = count[:n].sum(dim=1).to(torch.int64)
            row = torch.zeros(n + 1, dtype=torch.int64, device=dev)
            row[1:] = torch.cumsum(tot, 0)
            row32 = row.to(torch.int32)
            total = int(row[-1].item()) if n > 0 else 0
            neighbs = torch.zeros(max(total, 1), dtype=torch.int32, device=dev)
            rc = sim.lib.tsph_neighbor_lists(P, ctypes.byref(o.struct()), U, npc, nps, pin, uin,
                                             row32.data_ptr(), neighbs.data_ptr(), shift.data_ptr(),
                                             count.data_ptr(), sim.stream_ptr())
            _lib.check(rc, "tsph_neighbor_lists")
            part.neighbs = neighbs[:total]
            part.row_start = row32
            part.neighbs_shift = shift[:n]
            part.neighbs_count = count[:n]

    def neighbor_first(self, obj, nobj, i):
        p = self.parts[obj.uid]
        return int(p.row_start[i]) + int(p.neighbs_shift[i, nobj.uid])

    def neighbor_last(self, obj, nobj, i):
        p = self.parts[obj.uid]
        return self.neighbor_first(obj, nobj, i) + int(p.neighbs_count[i, nobj.uid])
