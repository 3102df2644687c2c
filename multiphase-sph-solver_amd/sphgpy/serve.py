"""Continuous-batching request loop: simulations of different sizes arrive over time, are admitted
against a device-memory budget, stay resident on the GPU (one context and one stream each) and
advance in step chunks that overlap on the device; each result is read back when its last chunk
has drained. Reports time-to-result per request and the aggregate particle-steps/s.

The reference runs one deck per process; this loop is the serving layer around the same step
library (sphg_step_async / sphg_poll)."""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from .capi import GpuSim, kl_field, setting_par
from .deck import Deck, column_collapse_2d, slope_2d, slope_pile_3d, soil_water_column_3d

SPHG_BUSY = -1
BYTES_PER_PARTICLE = {2: 1200, 3: 1700}  # resident state + 99-entry lists, rounded up


@dataclass
class Request:
    rid: int
    kind: str
    arrival: float                    # seconds after the stream starts
    steps: int
    make: Callable[[], Deck]          # deck factory
    kl: Optional[Tuple] = None        # (dist, mean, sd, rl, xi) for a random-field cohesion
    # filled in by the server
    admitted: float = float("nan")
    finished: float = float("nan")
    n: int = 0
    result: Optional[Dict[str, np.ndarray]] = None
    error: int = 0

    @property
    def ttr(self) -> float:
        return self.finished - self.arrival


def prepare(req: Request, deck: Optional[Deck] = None) -> Tuple[GpuSim, Deck]:
    """Context for one request: deck, initial state, optional K-L cohesion, step counter 1."""
    deck = deck if deck is not None else req.make()
    g = GpuSim(deck)
    g.load_state(setting_par(deck))
    g.upload_done()
    if req.kl is not None:
        dist, mean, sd, rl, xi = req.kl
        g.kl_synthesize(kl_field(deck, dist=dist, mean=mean, sd=sd, rl=rl), xi)
    g.set_step(1)
    return g, deck


RESULT_FIELDS = ("xp", "vx", "sig", "rho", "eps")


def run_alone(req: Request) -> Dict[str, np.ndarray]:
    g, _ = prepare(req)
    rc = g.step(req.steps)
    if rc:
        raise RuntimeError(f"request {req.rid}: error {rc}")
    out = {k: g.get(k) for k in RESULT_FIELDS}
    g.close()
    return out


@dataclass
class _Active:
    req: Request
    sim: GpuSim
    left: int
    bytes: int


class Server:
    def __init__(self, hbm_budget: float = 64e9, chunk: int = 50, max_active: int = 12):
        self.budget = hbm_budget
        self.chunk = chunk
        self.max_active = max_active
        self.used = 0
        self.log: List[str] = []

    def run(self, requests: List[Request]) -> Dict[str, float]:
        pending = sorted(requests, key=lambda r: r.arrival)
        active: List[_Active] = []
        done: List[Request] = []
        decks: Dict[int, Deck] = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        while pending or active:
            # admission: arrived requests in arrival order, while the memory budget and slots allow
            while pending and pending[0].arrival <= now() and len(active) < self.max_active:
                req = pending[0]
                if req.n == 0:
                    decks[req.rid] = req.make()
                    req.n = decks[req.rid].n
                need = req.n * BYTES_PER_PARTICLE[decks[req.rid].ndim]
                if self.used + need > self.budget and active:
                    break  # wait for memory to free up
                pending.pop(0)
                sim, deck = prepare(req, decks.pop(req.rid))
                req.n = deck.n
                nb = sim.device_bytes()
                self.used += nb
                req.admitted = now()
                active.append(_Active(req, sim, req.steps, nb))
            # progress: every drained context gets its next chunk, finished ones return results
            progressed = False
            for a in list(active):
                st = a.sim.L.lib.sphg_poll(a.sim.ctx)
                if st == SPHG_BUSY:
                    continue
                progressed = True
                if st != 0:
                    a.req.error = int(st)
                    a.left = 0
                if a.left > 0:
                    k = min(self.chunk, a.left)
                    rc = a.sim.L.lib.sphg_step_async(a.sim.ctx, k)
                    if rc:
                        a.req.error = int(rc)
                        a.left = 0
                    else:
                        a.left -= k
                    continue
                a.req.result = {k: a.sim.get(k) for k in RESULT_FIELDS}
                a.req.finished = now()
                a.sim.close()
                self.used -= a.bytes
                active.remove(a)
                done.append(a.req)
            if not progressed:
                time.sleep(0.0002)
        makespan = now()
        ttr = np.array([r.ttr for r in done])
        work = float(sum(r.n * r.steps for r in done if not r.error))
        return {"requests": len(done), "failed": int(sum(1 for r in done if r.error)),
                "errors": [(r.rid, r.kind, r.n, r.error) for r in done if r.error], "makespan_s": makespan,
                "particle_steps_per_s": work / makespan, "p50_ttr_s": float(np.percentile(ttr, 50)),
                "p99_ttr_s": float(np.percentile(ttr, 99)), "max_ttr_s": float(ttr.max())}


KINDS = ("column2d", "slope2d", "soilwater3d", "slopepile3d")


def request_stream(n: int, rate: float, seed: int = 0, steps: Tuple[int, int] = (500, 5000),
                   scale: float = 1.0, cycle: bool = False) -> List[Request]:
    """Poisson arrivals at `rate` per second of requests drawn by seed from four families:
    2D column collapse (5k-50k particles), 2D random-field slope (20k-80k), 3D soil-water column
    (100k-500k) and 3D slope with an embedded pile (0.5M-2M). `scale` shrinks the particle counts
    (tests); `cycle` takes the families in turn instead of at random."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        pick = int(rng.integers(0, len(KINDS)))
        kind = KINDS[rid % len(KINDS)] if cycle else KINDS[pick]
        st = int(rng.integers(steps[0], steps[1] + 1))
        s = int(rng.integers(0, 1 << 30))
        if kind == "column2d":
            target = rng.uniform(5e3, 5e4) * scale
            nx = max(4, int(np.sqrt(target / 2.0)))
            make = (lambda nx=nx, s=s: column_collapse_2d(nx=nx, ny=2 * nx, wall_h=2 * nx + 10,
                                                          floor_len=3 * nx, steps=10, seed=s))
            kl = None
        elif kind == "slope2d":
            target = rng.uniform(2e4, 8e4) * scale
            h = 10.0
            dx = float(np.sqrt(slope_2d(h, dx=0.1).n * 0.01 / target))
            make = (lambda h=h, dx=dx, s=s: slope_2d(h, dx=dx, steps=10, seed=s))
            kl = (2, 10000.0, 3000.0, (10.0, 2.0), rng.standard_normal(200))
        elif kind == "soilwater3d":
            target = rng.uniform(1e5, 5e5) * scale
            k = max(2, int(round((target / 5.0) ** (1.0 / 3.0))))  # soil 4k^3 + water k^3 + walls
            make = (lambda k=k, s=s: soil_water_column_3d(nx=2 * k, ny=k, nz_soil=2 * k, nz_water=max(1, k // 2),
                                                          steps=10, seed=s))
            kl = None
        else:
            target = rng.uniform(5e5, 2e6) * scale
            k = max(4, int(round((target / 8.6) ** (1.0 / 3.0))))  # 4k x k x 2k lattice, 1:2 cut, walls
            make = (lambda k=k, s=s: slope_pile_3d(nx=4 * k, ny=k, nz=2 * k, pile=(max(2, k // 5), max(2, k // 4)),
                                                   steps=10, seed=s, v0=(-0.3, 0.1, -0.05)))
            kl = None
        out.append(Request(rid, kind, t, st, make, kl))
    return out
