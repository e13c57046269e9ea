"""Serving loop: many KMC simulations resident on one GPU, continuously batched.

A request is one simulation (2D or 3D grid, temperature, seed) that runs a
fixed number of KMC events, each with its LS re-equilibration in the FP32 +
FP64 refinement form.  The server

  * admits arrived requests in arrival order while the estimated HBM footprint
    of the admitted set stays within the budget and fewer than max_active
    simulations are resident;
  * groups the resident simulations by grid shape: each group has one
    runner thread and HIP stream, which steps its simulations together in the
    batched native event loop (mgk_kmc_run_batch: one synchronisation per
    step for all rate totals, one for all selections, one LS solve with one
    launch per sweep for every member that had an event), a chunk of events
    at a time.  Small grids, whose launches are latency bound, share a group
    (up to 8 members); grids that fill the GPU alone get groups of one or
    two, so same-shape large requests still overlap on separate streams;
  * lets requests join and leave a running group between chunks: a request
    that arrives while others of its shape run is built on the group's stream
    and joins at the next chunk boundary, and a finished one leaves at the
    end of its chunk (continuous batching, not a queue in front of one
    simulation); groups of different shapes overlap on the device;
  * keeps each simulation's fields, modes and KMC state in HBM for its whole
    life; its random streams are private (numpy RandomState for waiting times,
    selection and redraws; Philox on the device for the initial modes), and a
    member of a batch runs the kernels of a simulation alone on its own
    buffers, so a request's result is bit-identical alone and in the mix.

There is no counterpart in the reference (its simulators run one at a time).
"""
import queue
import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor, wait, FIRST_COMPLETED
from dataclasses import dataclass, field

import ctypes
import numpy as np

from . import _lib
from . import device as dev
from .linear_elastic_simulator import LsSolver, MIXED, FP64, FP32

MU_UNIT = 70e9 / 2.6   # Pa: E = 70 GPa, nu = 0.3
E0_EV = 2.2            # barrier mean (eV): T = (kT/E0) * E0 / kB
KB = 8.617e-5
APPLIED_STRAIN = 0.005
SIZES_2D = ((64, 64), (128, 128), (256, 256), (512, 512))
SIZES_3D = ((64, 64, 64), (96, 96, 96), (128, 128, 128), (192, 192, 192))


@dataclass
class SimRequest:
    rid: int
    arrival: float          # seconds after the start of the stream
    shape: tuple
    kT: float               # kT / E0
    seed: int
    events: int = 100
    ls_precision: str = "mixed"
    M: int = 20


@dataclass
class SimResult:
    rid: int
    shape: tuple
    events: list
    eps_macro: np.ndarray
    sig_mean: np.ndarray
    time: float
    eps_p_sum: float        # FP64 sum of the eigenstrain field (a fingerprint of the flips)
    ls_iters: int
    footprint: int
    t_arrival: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.t_done - self.t_arrival


def request_stream(n, rate=4.0, seed=4, events=100, sizes=SIZES_2D + SIZES_3D, kT=(0.01, 0.08),
                   ls_precision="mixed"):
    """n requests with Poisson arrivals (mean rate `rate` per second), grid
    sizes uniform over `sizes`, kT/E0 ~ U[kT], per-request seeds, from seed."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps)
    idx = rng.integers(0, len(sizes), n)
    kts = rng.uniform(kT[0], kT[1], n)
    seeds = rng.integers(1, 2 ** 31 - 1, n)
    return [SimRequest(i, float(arrivals[i]), tuple(sizes[idx[i]]), float(kts[i]), int(seeds[i]), events,
                       ls_precision) for i in range(n)]


def _kmc_bytes(shape, M, eps_prec):
    d = len(shape)
    C = 6 if d == 3 else 3
    NV = int(np.prod(shape))
    n = ctypes.c_size_t()
    _lib.check(_lib.load().mgk_kmc_workspace_bytes(NV, ctypes.byref(n)), "mgk_kmc_workspace_bytes")
    es = 4 if eps_prec == FP32 else 8
    shadow = 4 * C * NV if eps_prec == MIXED else 0
    return NV * M * 8 * 3 + NV * 8 * 3 + C * NV * es + shadow + int(n.value)


def footprint_bytes(req):
    """Estimated HBM bytes of a request's resident state (LS fields and
    workspaces, modes, KMC state) plus transient upload buffers."""
    prec = {"mixed": MIXED, "fp64": FP64, "fp32": FP32}[req.ls_precision]
    NV = int(np.prod(req.shape))
    C = 6 if len(req.shape) == 3 else 3
    ls = LsSolver.footprint_bytes(req.shape, prec)
    kmc = _kmc_bytes(req.shape, req.M, prec)
    transient = NV * 8 * (C + 2)
    return int(1.05 * (ls + kmc + transient)) + (16 << 20)


def build_simulation(req):
    """The request's simulation with the serving recipe: mu ~ U[0.9, 1.1] *
    26.92 GPa, nu = 0.35, zero-mean Gaussian eigenstrain (std 1e-3 per stored
    component, drawn on the device from the request's seed: a 192^3 field is
    42M normals, most of a request's set-up time if drawn on the host), M
    modes, gamma0 0.14, pixel 0.7 nm, Gaussian barriers (2.2,
    0.45, clipped at 0) drawn on the device, jp 20, jt 0, cap 0.916, and a
    fixed macro strain eps_xx = APPLIED_STRAIN (strain rate 0: thermally
    activated relaxation under load; a driven strain rate would strain a
    small cold grid by ~10% per event, since waiting times grow as e^(Q/kT))."""
    from .aqs import ThermalSimulation
    from .kmc_simulator import KmcSimulation2D
    rng = np.random.default_rng(req.seed)
    shape = tuple(req.shape)
    d = len(shape)
    mu = MU_UNIT * rng.uniform(0.9, 1.1, shape)
    nu = np.full(shape, 0.35)
    E = 2.0 * mu * (1.0 + nu)
    C = 6 if d == 3 else 3
    T = req.kT * E0_EV / KB
    common = dict(pixel=0.7, barrier_generator="gaussian", barrier_kwargs=dict(mean=2.2, std=0.45, min_cutoff=0.0),
                  softening_cap=0.916, jp=20.0, jt=0.0, temperature=T, strain_rate=0.0,
                  ls_precision=req.ls_precision, rng=np.random.RandomState(req.seed),
                  selection_stream=np.random.RandomState(req.seed + 1), init_modes=("device", req.seed))
    if d == 3:
        sim = ThermalSimulation(*shape, req.M, 0.14, E, nu, **common)
    else:
        sim = KmcSimulation2D(*shape, req.M, 0.14, E, nu, **common)
    torch = dev.torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(req.seed)
    ep = torch.randn((C,) + shape, generator=gen, dtype=torch.float64, device="cuda") * 1e-3
    ep -= ep.mean(dim=tuple(range(1, d + 1)), keepdim=True)
    prec = sim.kmc.eps_p.dtype
    sim.kmc.eps_p.copy_(ep.to(prec))
    del ep
    if sim.kmc.eps_p32 is not None:
        sim.kmc.eps_p32.copy_(sim.kmc.eps_p.to(dev.torch.float32))
    sim.eps_macro[0, 0] = APPLIED_STRAIN
    return sim


_tls = threading.local()


def _worker_stream():
    s = getattr(_tls, "stream", None)
    if s is None:
        s = dev.torch.cuda.Stream()  # default priority
        _tls.stream = s
    return s


def run_request(req):
    """Run one request to completion on the calling thread's HIP stream."""
    torch = dev.torch
    s = _worker_stream()
    with torch.cuda.stream(s):
        sim = build_simulation(req)
        fp = sim.footprint_bytes()
        d = sim.ndim
        unit = np.zeros((d, d))
        unit[0, 0] = 1.0
        err = None
        try:
            sim.update_stresses()
            sim.kmc_events(1e30, unit, req.events)
        except (ValueError, FloatingPointError) as e:  # the simulators' solver-failure errors
            err = str(e)
        res = _result(req, sim, fp, err)
        s.synchronize()
        del sim
    return res


def _result(req, sim, fp, err=None):
    torch = dev.torch
    eps_p_sum = float(sim.kmc.eps_p.to(torch.float64).sum())
    res = SimResult(req.rid, tuple(req.shape), list(sim.kmc_event_history), sim.eps_macro.copy(),
                    sim.sig_mean().copy() if sim.ls_info is not None else None, sim.time, eps_p_sum,
                    sim.ls_iters_total, fp)
    if err is not None:
        res.extra["error"] = err
    return res


class _Member:
    __slots__ = ("req", "sim", "fp", "left", "rem", "seq", "unit")

    def __init__(self, req, sim, fp):
        self.req, self.sim, self.fp = req, sim, fp
        self.left = req.events
        self.rem = 1e30
        self.seq = 0
        d = sim.ndim
        self.unit = np.zeros((d, d))
        self.unit[0, 0] = 1.0


class _Group:
    """Resident simulations of one grid shape stepping together (at most cap)
    and their runner's inbox; n counts the members assigned and not yet
    reported done."""

    def __init__(self, cap):
        self.lock = threading.Lock()
        self.inbox = deque()
        self.running = False
        self.cap = cap
        self.n = 0


def batch_cap(shape, voxels=1 << 21, limit=None):
    """Members per batch for a grid shape: batching pays where a member's
    launches are latency bound (small grids); a grid whose sweeps fill the
    GPU on their own gains nothing from it and would only step in lockstep
    with its peers, so up to `voxels` voxels per batch step, at least one."""
    limit = _lib.MGK_LS_MAX_BATCH if limit is None else limit
    nv = int(np.prod(shape))
    return max(1, min(limit, voxels // nv))


class SimServer:
    """Admission by HBM footprint; continuous batching per grid shape (groups
    of up to batch_cap(shape) members, each group with its own runner thread
    and HIP stream, chunk_events events per member between joins).
    batch=False runs every request alone on its own worker thread and stream
    instead (no batching; kept for comparison)."""

    def __init__(self, hbm_budget=None, max_active=16, batch=True, chunk_events=10, batch_voxels=1 << 21):
        dev.require_gpu()
        if hbm_budget is None:
            free, _total = dev.torch.cuda.mem_get_info()
            hbm_budget = int(0.85 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.batch = bool(batch)
        self.chunk_events = int(chunk_events)
        self.batch_voxels = int(batch_voxels)
        self.peak_used = 0
        self.peak_active = 0
        self.peak_batch = 0
        self.batched_steps = 0

    # ----- batched runner ------------------------------------------------------
    def _runner(self, grp, done_q):
        from .kmc_engine import kmc_events_batch
        torch = dev.torch
        s = torch.cuda.Stream()  # default priority
        members = []
        with torch.cuda.stream(s):
            while True:
                new = []
                with grp.lock:
                    while grp.inbox and len(members) + len(new) < grp.cap:
                        new.append(grp.inbox.popleft())
                    if not members and not new:
                        grp.running = False
                        return
                for req, fp in new:
                    try:
                        sim = build_simulation(req)
                        sim.update_stresses()
                        members.append(_Member(req, sim, fp))
                    except (ValueError, FloatingPointError) as e:
                        done_q.put((req, fp, SimResult(req.rid, tuple(req.shape), [], None, None, 0.0, 0.0, 0, fp,
                                                       extra={"error": str(e)})))
                if not members:
                    continue
                self.peak_batch = max(self.peak_batch, len(members))
                k = [min(self.chunk_events, m.left) for m in members]
                outs = kmc_events_batch([m.sim for m in members], [m.rem for m in members],
                                        [m.unit for m in members], k, seqs=[m.seq for m in members])
                self.batched_steps += max(k)
                keep = []
                for m, (n, rem, reason, seq, err) in zip(members, outs):
                    m.left -= n
                    m.rem, m.seq = rem, seq
                    if reason == "count" and m.left > 0:
                        keep.append(m)
                        continue
                    res = _result(m.req, m.sim, m.fp, None if err is None else str(err))
                    s.synchronize()
                    m.sim = None
                    done_q.put((m.req, m.fp, res))
                members = keep

    def _serve_batched(self, pending, t0, on_done, results, time_scale):
        done_q = queue.Queue()
        groups = {}
        queue_ = deque()
        used = 0
        active = 0
        while pending or queue_ or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival * time_scale <= now:
                queue_.append(pending.popleft())
            while queue_ and active < self.max_active:
                fp = footprint_bytes(queue_[0])
                if used + fp > self.hbm_budget:
                    break
                r = queue_.popleft()
                used += fp
                active += 1
                r._t_admit = time.perf_counter() - t0
                # a group of this shape with room (a running one first), else a new group
                gl = groups.setdefault(tuple(r.shape), [])
                free = [g for g in gl if g.n < g.cap]
                free.sort(key=lambda g: not g.running)
                if free:
                    g = free[0]
                else:
                    g = _Group(batch_cap(r.shape, self.batch_voxels))
                    gl.append(g)
                g.n += 1
                r._group = g
                with g.lock:
                    g.inbox.append((r, fp))
                    start = not g.running
                    g.running = True
                if start:
                    threading.Thread(target=self._runner, args=(g, done_q), daemon=True).start()
            self.peak_used = max(self.peak_used, used)
            self.peak_active = max(self.peak_active, active)
            timeout = None
            if pending:
                timeout = max(0.0, pending[0].arrival * time_scale - (time.perf_counter() - t0))
            if active:
                try:
                    item = done_q.get(timeout=timeout)
                except queue.Empty:
                    continue
                while True:
                    r, fp, res = item
                    used -= fp
                    active -= 1
                    r._group.n -= 1
                    res.t_arrival = r.arrival * time_scale
                    res.t_admit = r._t_admit
                    res.t_done = time.perf_counter() - t0
                    results.append(res)
                    if on_done is not None:
                        on_done(res)
                    try:
                        item = done_q.get_nowait()
                    except queue.Empty:
                        break
            elif timeout:
                time.sleep(timeout)

    # ----- one worker per request (no batching) --------------------------------
    def _serve_unbatched(self, pending, t0, on_done, results, time_scale):
        queue_ = deque()
        active = {}
        used = 0
        with ThreadPoolExecutor(max_workers=self.max_active) as pool:
            while pending or queue_ or active:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival * time_scale <= now:
                    queue_.append(pending.popleft())
                while queue_ and len(active) < self.max_active:
                    fp = footprint_bytes(queue_[0])
                    if used + fp > self.hbm_budget:
                        break
                    r = queue_.popleft()
                    used += fp
                    fut = pool.submit(run_request, r)
                    active[fut] = (r, fp, time.perf_counter() - t0)
                self.peak_used = max(self.peak_used, used)
                self.peak_active = max(self.peak_active, len(active))
                timeout = None
                if pending:
                    timeout = max(0.0, pending[0].arrival * time_scale - (time.perf_counter() - t0))
                if active:
                    done, _ = wait(list(active), timeout=timeout, return_when=FIRST_COMPLETED)
                    for fut in done:
                        r, fp, t_admit = active.pop(fut)
                        used -= fp
                        res = fut.result()
                        res.t_arrival = r.arrival * time_scale
                        res.t_admit = t_admit
                        res.t_done = time.perf_counter() - t0
                        results.append(res)
                        if on_done is not None:
                            on_done(res)
                elif timeout:
                    time.sleep(timeout)

    def serve(self, requests, time_scale=1.0, on_done=None):
        """Replay `requests` against the wall clock (arrival * time_scale) and
        return their SimResults in completion order."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        for r in requests:
            if footprint_bytes(r) > self.hbm_budget:
                raise ValueError(f"request {r.rid} ({r.shape}) needs more than the HBM budget")
        results = []
        t0 = time.perf_counter()
        if self.batch:
            self._serve_batched(pending, t0, on_done, results, time_scale)
        else:
            self._serve_unbatched(pending, t0, on_done, results, time_scale)
        self.wall = time.perf_counter() - t0
        return results


def summarize(results, wall):
    lat = np.array([r.latency for r in results])
    ev = sum(len(r.events) for r in results)
    return {"requests": len(results), "failed": sum(1 for r in results if "error" in r.extra),
            "wall_s": wall, "requests_per_s": len(results) / wall,
            "kmc_events_per_s": ev / wall, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "latency_mean_s": float(lat.mean()),
            "latency_max_s": float(lat.max())}


def main(argv=None):
    """python -m mgkmc_amd.serving [--requests N] [--rate R] [--events E] [--slots S]:
    replay a Poisson request stream and print one JSON summary line."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--events", type=int, default=100)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--no-batch", action="store_true", help="one worker per request, no batching")
    ap.add_argument("--chunk", type=int, default=10, help="events per member between joins")
    a = ap.parse_args(argv)
    reqs = request_stream(a.requests, a.rate, a.seed, a.events)
    srv = SimServer(max_active=a.slots, batch=not a.no_batch, chunk_events=a.chunk)
    done = [0]

    def progress(res):
        done[0] += 1
        if done[0] % 16 == 0:
            print(f"[serving] {done[0]}/{len(reqs)} done", flush=True)
    res = srv.serve(reqs, on_done=progress)
    out = summarize(res, srv.wall)
    out.update(slots=a.slots, rate=a.rate, events_per_request=a.events, peak_active=srv.peak_active,
               peak_hbm_bytes=srv.peak_used, hbm_budget=srv.hbm_budget, batched=srv.batch,
               peak_batch=srv.peak_batch, chunk_events=srv.chunk_events)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
