"""Batched solves: many ``FDTD`` problems advanced together on one GPU.

Each member keeps its own grid size, wavelength, materials, sources, Tn and
convergence test; one kernel launch per half step covers all members, and
every member's solve loop (emopt/fdtd.py:994-1194) runs on the device, so the
host only replays a captured hipGraph of ``steps_per_graph`` steps and reads
the members' states between replays.  Members may be added while others
run (continuous batching).  A member's fields agree with those of
``solve_forward`` / ``solve_adjoint`` on the same problem to roundoff (same
n0 and step counts), and do not depend on the other members.

    b = Batch()
    ids = [b.add(sim) for sim in sims]          # forward solves
    b.add(other, adjoint=True)                  # an adjoint solve in the same batch
    b.run()                                     # until all have finished
    sims[0].get_field("Ez")
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import SolveInfo

__all__ = ["Batch", "GROUP_BYTES", "LANES"]

# Members are stepped in groups whose hot state (~80 B/cell) stays within
# GROUP_BYTES, and LANES groups are stepped at once (one stream each), so the
# groups in flight together -- about MI355X's 256 MB Infinity Cache -- let each
# group's E pass re-read from cache what its H pass just wrote, while one
# lane's kernels fill the CUs that the other's leave idle at their tails.
GROUP_BYTES = 130 * 10**6
LANES = 2


class Batch:
    def __init__(self, device=0, steps_per_graph=64, zc=2, blocks_per_cu=16,
                 group_bytes=GROUP_BYTES, rotate_groups=False, use_graph=True, lanes=LANES):
        self._lib = _lib.load()
        self.device = device
        h = ctypes.c_void_p()
        self._lib.emf_batch_new(device, ctypes.byref(h))
        self._h = h.value
        self._members = {}          # id -> (sim, slot, postprocess)
        self.finished = {}          # id -> solve info dict of finished members
        self.set_tuning(steps_per_graph, zc, blocks_per_cu)
        self.set_grouping(group_bytes, rotate_groups)
        self.set_graph(use_graph)
        self.set_lanes(lanes)

    def set_lanes(self, lanes):
        """Step this many groups at once (1..4), each on its own stream."""
        self._lib.emf_batch_set_lanes(self._h, int(lanes))
        self.lanes = int(lanes)

    def set_graph(self, use_graph):
        """True (default): one graph launch per replay.  False: the replay's
        kernels are launched one by one (for per-kernel tracing)."""
        self._lib.emf_batch_set_graph(self._h, 1 if use_graph else 0)
        self.use_graph = bool(use_graph)

    def set_grouping(self, group_bytes, rotate=False):
        """Step members in groups whose hot state fits `group_bytes`, one
        group per graph replay: the first group until its members finish
        (rotate=False), or the groups in turns (rotate=True).  0 steps all
        members together."""
        self._lib.emf_batch_set_grouping(self._h, int(group_bytes), 1 if rotate else 0)
        self.group_bytes = group_bytes
        self.rotate_groups = bool(rotate)

    def set_tuning(self, steps_per_graph, zc, blocks_per_cu):
        self._lib.emf_batch_set_tuning(self._h, int(steps_per_graph), int(zc), int(blocks_per_cu))
        self.tuning = dict(steps_per_graph=steps_per_graph, zc=zc, blocks_per_cu=blocks_per_cu)

    def add(self, sim, adjoint=False, postprocess=True):
        """Queue the forward (or adjoint) solve of `sim`; returns the member
        id.  With postprocess, a finished forward solve updates the saved
        fields and source power as ``solve_forward`` does."""
        if not sim._mats_ready:
            sim.build()
        sim._register(sim._adj_sources if adjoint else sim._sources)
        slot = 1 if adjoint else 0
        n0 = -1 if sim.fixed_n0 is None else int(sim.fixed_n0)
        mid = ctypes.c_int()
        self._lib.emf_batch_add(self._h, sim._h, slot, float(sim._rtol), float(sim.Nmax), sim.Tn,
                                n0, ctypes.byref(mid))
        self._members[mid.value] = (sim, slot, postprocess)
        return mid.value

    def run(self, until="all", max_graphs=1 << 30):
        """Replay the step graph.  until = "all" (every member finished),
        "any" (return at the first finished member) or "graphs" (exactly
        max_graphs replays).  Finished members are detached; returns their
        ids."""
        u = {"graphs": 0, "any": 1, "all": 2}[until]
        nf = ctypes.c_int()
        self._lib.emf_batch_run(self._h, int(max_graphs), u, ctypes.byref(nf))
        return self._collect()

    def _collect(self):
        done = []
        for mid, (sim, slot, post) in list(self._members.items()):
            info = SolveInfo()
            if self._lib.emf_batch_poll(self._h, mid, ctypes.byref(info)) != 1:
                continue
            self._lib.emf_batch_remove(self._h, mid)
            del self._members[mid]
            sim._solved(slot, info)
            if slot == 0:
                if post:
                    sim._forward_done()
            else:
                sim.adjoint_count += 1
            self.finished[mid] = dict(sim.last_solve)
            done.append(mid)
        return done

    def free_bytes(self):
        """Free device memory of the batch's GPU (hipMemGetInfo)."""
        f, t = ctypes.c_int64(), ctypes.c_int64()
        self._lib.emf_mem_info(self.device, ctypes.byref(f), ctypes.byref(t))
        return int(f.value)

    @property
    def running(self):
        r, f = ctypes.c_int(), ctypes.c_int()
        self._lib.emf_batch_count(self._h, ctypes.byref(r), ctypes.byref(f))
        return r.value

    def remove(self, mid):
        """Detach a member without waiting for it to finish."""
        self._lib.emf_batch_remove(self._h, mid)
        self._members.pop(mid, None)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.emf_batch_delete(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
