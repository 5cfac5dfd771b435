"""Serving loop policy on the CPU (a stand-in batch, no GPU): admission by
the HBM budget (never exceeded), first-come first-served order, requests
larger than the whole budget rejected, every admitted request completed, and
latency statistics computed from the recorded times."""
import numpy as np

from emopt_amd.serve import Request, Server, state_bytes


class FakeBatch:
    """Finishes each member after `life` polls; records the membership."""

    def __init__(self, life):
        self.life = life
        self.members = {}
        self.finished = {}
        self.order = []
        self.next_id = 0
        self.max_members = 0

    def add(self, sim, adjoint=False, postprocess=True):
        mid = self.next_id
        self.next_id += 1
        self.members[mid] = self.life[sim]
        self.order.append(sim)
        self.max_members = max(self.max_members, len(self.members))
        return mid

    def run(self, until="any", max_graphs=1):
        done = []
        for mid in list(self.members):
            self.members[mid] -= 1
            if self.members[mid] <= 0:
                del self.members[mid]
                self.finished[mid] = dict(steps=10)
                done.append(mid)
        return done

    def close(self):
        pass


def test_state_bytes():
    # 10x10x10 grid, no CPML: 220 B per cell with one result slot
    assert state_bytes(10, 10, 10, [0] * 6) == 1000 * 220
    assert state_bytes(10, 10, 10, [0] * 6, slots=2) == 1000 * 316
    # psi: 2 components x (E and H) x 8 B per slab cell
    assert state_bytes(10, 10, 10, [1, 0, 0, 0, 0, 0]) == 1000 * 220 + 2 * 100 * 2 * 8
    # with the wavelength: + 16 B/cell of sampler staging, + 64 B per step of
    # the step-factor table (500 wavelength/d steps and the capture periods)
    base = state_bytes(10, 10, 10, [0] * 6)
    full = state_bytes(10, 10, 10, [0] * 6, wavelength=1.5, d=0.05)
    assert full - base >= 1000 * 16 + 64 * 500 * 30
    assert full - base < 1000 * 16 + 64 * 510 * 30 + 10000


class FreeMemBatch(FakeBatch):
    """A stand-in batch whose device has `total` bytes: each running member
    holds its request's bytes until it finishes."""

    def __init__(self, life, sizes, total):
        super().__init__(life)
        self.sizes, self.total, self.held = sizes, total, {}
        self.peak = 0

    def add(self, sim, adjoint=False, postprocess=True):
        mid = super().add(sim, adjoint, postprocess)
        self.held[mid] = self.sizes[sim]
        self.peak = max(self.peak, sum(self.held.values()))
        return mid

    def run(self, until="any", max_graphs=1):
        done = super().run(until, max_graphs)
        for mid in done:
            self.held.pop(mid)
        return done

    def free_bytes(self):
        return self.total - sum(self.held.values())


def test_admission_waits_for_free_device_memory():
    """With a budget far above the device, admission follows the device's
    free memory (less the reserve): the members held at once never exceed
    it, requests still complete in arrival order."""
    rng = np.random.default_rng(3)
    sizes = {q: int(rng.integers(20, 60)) for q in range(30)}
    life = {q: int(rng.integers(1, 5)) for q in range(30)}
    fb = FreeMemBatch(life, sizes, total=130)
    srv = Server(hbm_budget=1e9, prep_workers=2, batch=fb, max_running=8, hbm_reserve=10)
    for q in range(30):
        srv.submit(Request(q, (lambda q=q: q), sizes[q], cells=1))
    srv.drain()
    srv.close()
    assert fb.peak <= 120
    assert fb.order == list(range(30))
    assert srv.stats()["completed"] == 30
    assert srv.waits_on_free > 0


def test_admission_budget_and_order():
    rng = np.random.default_rng(0)
    sizes = rng.integers(10, 60, size=40)
    life = {q: int(rng.integers(1, 6)) for q in range(40)}
    fb = FakeBatch(life)
    budget = 150
    used_trace = []
    srv = Server(hbm_budget=budget, prep_workers=2, batch=fb)
    reqs = [Request(q, (lambda q=q: q), int(sizes[q]), cells=1) for q in range(40)]
    reqs.append(Request(99, lambda: 99, budget + 1, cells=1))    # can never fit
    for r in reqs:
        srv.submit(r)
    while not srv.idle:
        srv.poll()
        used_trace.append(srv.used)
        assert srv.used <= budget
    srv.close()
    assert fb.order == list(range(40))                 # admitted in arrival order
    st = srv.stats()
    assert st["completed"] == 40 and st["failed"] == 1
    assert max(used_trace) <= budget and srv.used == 0
    assert st["p99_s"] >= st["p50_s"] >= 0


def test_max_running_bounds_concurrency():
    rng = np.random.default_rng(1)
    life = {q: int(rng.integers(1, 4)) for q in range(20)}
    fb = FakeBatch(life)
    srv = Server(hbm_budget=1e9, prep_workers=3, batch=fb, max_running=2)
    for q in range(20):
        srv.submit(Request(q, (lambda q=q: q), 10, cells=1))
    srv.drain()
    srv.close()
    assert fb.max_members <= 2
    assert fb.order == list(range(20))
    assert srv.stats()["completed"] == 20
