"""A served stream of mass-matrix solves with continuous batching.

Requests arrive over time; each asks for A x = b on a cubed-sphere W0 space (edge cells,
levels, element order) with a number of right-hand sides and a tolerance.  The server admits a
request when its device footprint fits the HBM budget: it builds the request's operators and
vectors on the device and adds it to a running batch (lfric_batch_*), so requests that arrive
while others are being solved join them between rounds and share every kernel launch.
Finished requests leave the batch and free their memory at once, which lets waiting requests
in.

Admission policy: arrival order, with backfill - a request that does not fit waits, and later
requests that do fit may go ahead of it, until it has waited `starve_s` seconds; from then on
nothing else is admitted until it fits.

Requests are generated on the device (counter-based generator) as a stand-in for an upload of
the operator; that set-up is part of the measured latency."""
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .batch import Batch, prepare_batch_copies
from .fem import FieldVector, FunctionSpace, LmaOperator


@dataclass
class Request:
    rid: int
    t_arrive: float          # seconds from the start of the stream
    n: int                   # cubed-sphere edge cells
    nl: int                  # levels
    order: int               # W0 element order (0: 8 dofs per cell, 1: 27)
    nrhs: int                # right-hand sides asked for
    tol: float               # relative residual target
    seed: int
    # filled by the server
    t_admit: float = math.nan
    t_done: float = math.nan
    iterations: int = 0
    outer: int = 0
    true_rel_res: float = math.nan
    status: int = -99
    footprint: int = 0


def padded_rhs(nrhs):
    """The batch takes 1, 2, 4 or 8 right-hand sides; extra columns are zero."""
    for p in (1, 2, 4, 8):
        if nrhs <= p:
            return p
    raise ValueError(nrhs)


def footprint_bytes(n, nl, order, nrhs):
    """Device bytes a request holds while admitted: FP64 and FP32 operators (level-aligned
    layout with 64-slot colour padding), the FP32 operator's batch copy (row-gathered for
    order 0; for order 1 the symmetric-packed triangles, one RHS's row sums and the batch's
    row sums of every RHS), x and b in FP64, the batch's four FP32 vectors and FP64 residual,
    the Jacobi diagonal and the space's maps."""
    nc = 6 * n * n
    ndf = 8 if order == 0 else 27
    undf = (nc + 2) * (nl + 1) if order == 0 else (nc + 12 * n * n + nc + 2) * (2 * nl + 1)
    nelem = nc * (nl + 1) + 4 * 64
    ent32 = -(-ndf * ndf // 4) * 4
    ent64 = -(-ndf * ndf // 2) * 2
    ops = nelem * (ent32 * 4 + ent64 * 8)
    r = padded_rhs(nrhs)
    if order == 0:
        copy = (undf + 64) * 2 * 4 * 8 * 4
    else:
        units = nc * nl
        copy = (units + 64) * 380 * 4 + units * 27 * 4 * (1 + r) + undf * 16
    vecs = undf * r * (8 + 8 + 4 * 4 + 8) + undf * 4 * 2
    maps = nc * ndf * 4 * 4
    return ops + copy + vecs + maps


def poisson_stream(rate, window_s, seed=0, meshes=(24, 48, 96, 192), weights=(0.4, 0.3, 0.2, 0.1),
                   levels=(30, 70), orders=(0, 1), nrhs_max=8, tols=(1e-6, 1e-8)):
    """Seeded Poisson arrivals with the request mix of the served configuration."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= window_s:
            return out
        out.append(Request(len(out), t, int(rng.choice(meshes, p=weights)), int(rng.choice(levels)),
                           int(rng.choice(orders)), int(rng.integers(1, nrhs_max + 1)), float(rng.choice(tols)),
                           int(rng.integers(0, 2**31))))


def admissions(waiting, used, budget, now, starve_s):
    """The requests of `waiting` (in arrival order) to admit at time `now` with `used` of
    `budget` bytes taken: arrival order with backfill - a request that does not fit is passed
    over by later ones that do - until the head of the queue has waited more than `starve_s`;
    then nothing behind a request that does not fit is admitted."""
    out = []
    head_starved = bool(waiting) and now - waiting[0].t_arrive > starve_s
    for r in waiting:
        if used + r.footprint <= budget:
            out.append(r)
            used += r.footprint
        elif head_starved:
            break
    return out


class _Live:
    """Device objects of one admitted request."""

    def __init__(self, req):
        self.req = req
        fs = FunctionSpace.cubedsphere(req.n, req.nl, _lib.W0, req.order)
        self.fs = fs
        self.op64 = LmaOperator(fs).generate_spd(req.seed, scaled=True)
        self.op32 = prepare_batch_copies(self.op64.convert(_lib.R32))
        self.diag32 = self.op32.mm_diagonal(FieldVector(fs, _lib.R32))
        r = padded_rhs(req.nrhs)
        self.nrhs = r
        self.b = FieldVector(fs, _lib.R64, r).fill_uniform(req.seed, 3)
        if r != req.nrhs:  # zero the padding columns
            B = self.b.export_field().reshape(fs.undf, r)
            B[:, req.nrhs:] = 0.0
            self.b.import_field(B.ravel())
        self.x = FieldVector(fs, _lib.R64, r)
        self.x.set_scalar(0.0)


@dataclass
class ServeReport:
    requests: list
    wall_s: float
    completed: int
    unfinished: int
    rejected: int
    solves_per_s: float
    rhs_per_s: float
    dofs_per_s: float
    p50_ms: float
    p99_ms: float
    queue_p50_ms: float
    queue_p99_ms: float
    peak_bytes: int
    extra: dict = field(default_factory=dict)


class SolveServer:
    def __init__(self, hbm_budget_bytes, starve_s=2.0, check_every=8, inner_r_tol=1e-4, max_iter=4000):
        self.budget = int(hbm_budget_bytes)
        self.starve_s = starve_s
        self.opts = dict(check_every=check_every, inner_r_tol=inner_r_tol, max_iter=max_iter)
        self.batches = {}   # tolerance -> Batch (the tolerance is a batch option)

    def _batch(self, tol):
        if tol not in self.batches:
            self.batches[tol] = Batch(r_tol=tol, a_tol=0.0, **self.opts)
        return self.batches[tol]

    def serve(self, requests, wall_limit_s=60.0, drain_limit_s=None, clock=time.perf_counter):
        """Serve `requests` (sorted by arrival) in real time; stop admitting when the wall
        limit is reached, finish the admitted requests until drain_limit_s (default twice the
        wall limit), and report what finished."""
        drain_limit_s = 2 * wall_limit_s if drain_limit_s is None else drain_limit_s
        pending = list(requests)
        waiting, live = [], {}
        used = peak = 0
        rejected = 0
        t0 = clock()
        while True:
            now = clock() - t0
            while pending and pending[0].t_arrive <= now:
                r = pending.pop(0)
                r.footprint = footprint_bytes(r.n, r.nl, r.order, r.nrhs)
                if r.footprint > self.budget:
                    r.status = -100  # can never fit
                    rejected += 1
                    continue
                waiting.append(r)
            # admission: arrival order with backfill, unless the head has starved
            if now < wall_limit_s:
                for r in admissions(waiting, used, self.budget, now, self.starve_s):
                    waiting.remove(r)
                    r.t_admit = clock() - t0
                    lv = _Live(r)
                    bt = self._batch(r.tol)
                    pid = bt.add(lv.op64, lv.op32, lv.diag32, lv.x, lv.b)
                    live[(r.tol, pid)] = lv
                    used += r.footprint
                    peak = max(peak, used)
            if not live:
                if not pending and not waiting:
                    break
                if now >= wall_limit_s:
                    break
                if not waiting and pending:  # idle until the next arrival
                    time.sleep(max(0.0, min(pending[0].t_arrive - now, 0.01)))
                continue
            for tol, bt in self.batches.items():
                if not any(k[0] == tol for k in live):
                    continue
                bt.run(1)
                for key in [k for k in live if k[0] == tol]:
                    pid = key[1]
                    if bt.finished(pid):
                        lv = live.pop(key)
                        res = bt.result(pid)[: lv.req.nrhs]
                        r = lv.req
                        r.t_done = clock() - t0
                        r.iterations = max(x.iterations for x in res)
                        r.outer = res[0].outer_iterations
                        r.true_rel_res = max(x.true_rel_res for x in res)
                        r.status = max(x.status for x in res)
                        bt.remove(pid)
                        used -= r.footprint
                        del lv
            if (now >= wall_limit_s and not live) or now >= drain_limit_s:
                break
        wall = clock() - t0
        done = [r for r in requests if not math.isnan(r.t_done)]
        lat = np.array([(r.t_done - r.t_arrive) * 1e3 for r in done]) if done else np.array([math.nan])
        qd = np.array([(r.t_admit - r.t_arrive) * 1e3 for r in done]) if done else np.array([math.nan])
        rhs = sum(r.nrhs for r in done)
        dofs = sum(footprint_undf(r) * r.nrhs for r in done)
        return ServeReport(requests, wall, len(done), len(requests) - len(done) - rejected, rejected,
                           len(done) / wall, rhs / wall, dofs / wall, float(np.percentile(lat, 50)),
                           float(np.percentile(lat, 99)), float(np.percentile(qd, 50)), float(np.percentile(qd, 99)),
                           peak)

    def close(self):
        for bt in self.batches.values():
            bt.close()
        self.batches = {}


def footprint_undf(r):
    nc = 6 * r.n * r.n
    return (nc + 2) * (r.nl + 1) if r.order == 0 else (nc + 12 * r.n * r.n + nc + 2) * (2 * r.nl + 1)
