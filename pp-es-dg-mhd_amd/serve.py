"""Continuous-batching request loop for PP-ES DG solves on one GPU.

Requests (problem, mesh, degree, end time) arrive over time. Each admitted
request becomes an instance of one `ppesdg.Batch`: its state stays resident
in HBM from admission to completion, and every round advances all admitted
instances together by one RK step (one launch per kernel class and degree,
replayed from a captured hipGraph while the membership is unchanged). A
request is admitted as soon as the device memory it needs fits in the budget
(and an active slot is free), so requests that arrive while others run join
the running batch; a finished instance leaves the batch and frees its memory
before the next round's admissions.

The reference runs one problem per process (drive.f90:32-343, main.f90);
this loop has no counterpart there. Each instance's result is the same as
the instance run alone (the kernels never mix instances' data).
"""
import collections
import time

import ppesdg
import workloads as W


class Server:
    def __init__(self, budget_bytes=200e9, max_active=512, rk=3, use_graph=True, keep_results=False):
        self.batch = ppesdg.Batch(rk)
        self.rk = rk
        self.budget = float(budget_bytes)
        self.max_active = max_active
        self.use_graph = use_graph
        self.keep_results = keep_results
        self.pending = collections.deque()
        self.active = {}
        self.used = 0
        self.high_water = 0        # largest device bytes held by admitted instances at once
        self.admitted = 0
        self.rejected = []
        self.rounds = 0

    def busy(self):
        return bool(self.pending or self.active)

    def submit(self, req, now):
        r = dict(req)
        r["t_submit"] = now
        r["bytes"] = ppesdg.instance_bytes(r["nx"], r["ny"], r["p"], self.rk)
        if r["bytes"] > self.budget:
            self.rejected.append(r)
            return
        self.pending.append(r)

    def _admit(self, now):
        while self.pending and len(self.active) < self.max_active:
            r = self.pending[0]
            if self.used + r["bytes"] > self.budget:
                break
            self.pending.popleft()
            dom, g, bc, bcs, cfl, oedg = W.PROBLEMS[r["prob"]]
            i = self.batch.add(r["nx"], r["ny"], r["p"], dom, g, bc, bcs)
            self.batch.init_problem(i, r["prob"])
            self.batch.start(i, r.get("t_final", -1.0), r.get("max_steps", 0), cfl, oedg, True)
            r["t_admit"] = now
            self.active[i] = r
            self.used += r["bytes"]
            self.admitted += 1
            self.high_water = max(self.high_water, self.used)

    def round(self, clock):
        """Admit what fits, advance every admitted instance by one RK step,
        retire the instances that finished. Returns the finished requests."""
        self._admit(clock())
        if not self.active:
            return []
        self.batch.step(1, self.use_graph)
        self.rounds += 1
        finished = []
        for i, r in list(self.active.items()):
            st = self.batch.status(i)
            if st["state"] in ("done", "failed"):
                r["t_done"] = clock()
                r["stats"] = st
                r["positivity"] = self.batch.positivity(i)  # (min rho, min p) over all its stages
                if self.keep_results:
                    r["result"] = self.batch.download_state(i, 0)
                self.batch.remove(i)
                self.used -= r["bytes"]
                del self.active[i]
                finished.append(r)
        return finished

    def close(self):
        self.batch.close()


def run_stream(requests, budget_bytes=200e9, use_graph=True, keep_results=False, wall_cap=None, progress=None,
               report=None):
    """Serve `requests` (sorted by `arrival`, in seconds) in real time.
    Returns the finished requests with t_submit / t_admit / t_done (seconds
    since the stream started), the wall time and the rejected requests. With
    `wall_cap` (seconds) the loop stops there; `report`, a dict, then also
    receives the requests still running (with their progress so far), still
    queued and never submitted, the admission count and the budget's
    high-water mark. `progress(msg)` is called about every 30 s."""
    srv = Server(budget_bytes, use_graph=use_graph, keep_results=keep_results)
    t0 = time.perf_counter()

    def clock():
        return time.perf_counter() - t0

    done = []
    k = 0
    last = 0.0
    while k < len(requests) or srv.busy():
        now = clock()
        if wall_cap is not None and now >= wall_cap:
            break
        while k < len(requests) and requests[k]["arrival"] <= now:
            srv.submit(requests[k], requests[k]["arrival"])
            k += 1
        if not srv.busy():
            time.sleep(max(0.0, min(requests[k]["arrival"] - now, (wall_cap or 1e30) - now)))
            continue
        done += srv.round(clock)
        if progress is not None and clock() - last >= 30.0:
            last = clock()
            progress(f"serving t={last:.0f}s: {len(done)} done, {len(srv.active)} running, {len(srv.pending)} queued, "
                     f"{srv.used / 1e9:.1f} GB held")
    wall = clock()
    if report is not None:
        running = []
        for i, r in srv.active.items():
            q = dict(r)
            q["stats"] = srv.batch.status(i)
            running.append(q)
        report.update(running=running, queued=list(srv.pending), unsubmitted=requests[k:],
                      admitted=srv.admitted, high_water_bytes=srv.high_water, rounds=srv.rounds)
    srv.close()
    return done, wall, srv.rejected


def run_sequential(requests, wall_cap=None, chunk=50, progress=None):
    """The plain path on the same stream: one request at a time through the
    single-instance solver, first come first served. Returns per-request
    (start, finish) times on the stream's clock and the busy time. With
    `wall_cap`, the stream's clock stops there: the request then running
    reports its progress (stats, no t_done) and later ones are not started.
    `progress(msg)` is called about every 30 s."""
    out = []
    free_at = 0.0
    busy = 0.0
    last = time.perf_counter()
    for r in requests:
        start = max(r["arrival"], free_at)
        if wall_cap is not None and start >= wall_cap:
            break
        dom, g, bc, bcs, cfl, oedg = W.PROBLEMS[r["prob"]]
        t_start = time.perf_counter()
        s = ppesdg.Solver(r["nx"], r["ny"], r["p"], dom, g, bc, bcs)
        s.init_problem(r["prob"])
        if wall_cap is None:
            st = s.solve(r.get("t_final", -1.0), r.get("max_steps", 0), cfl, oedg, True)
            finished = True
        else:
            # the same solve in chunks of steps, to stop at the cap
            s.solve_begin(True)
            steps = 0
            while True:
                st = s.advance(r.get("t_final", -1.0), chunk, cfl, oedg)
                # fewer steps than asked: the solve reached its end (or failed)
                finished = st["status"] != 0 or st["steps"] < steps + chunk
                steps = st["steps"]
                if progress is not None and time.perf_counter() - last >= 30.0:
                    last = time.perf_counter()
                    progress(f"sequential t={start + (last - t_start):.0f}s: {len(out)} done, request "
                             f"{len(out)} at step {steps}")
                if finished or start + (time.perf_counter() - t_start) >= wall_cap:
                    break
        s.close()
        d = time.perf_counter() - t_start
        free_at = start + d
        busy += d
        q = dict(r)
        q.update(stats=st, solve_s=d)
        if finished:
            q["t_done"] = free_at
        out.append(q)
    return out, busy


def percentiles(latencies, qs=(50, 99)):
    import numpy as np
    a = np.asarray(latencies, dtype=float)
    return {f"p{q}": float(np.percentile(a, q)) for q in qs}


def work(r):
    """DOF-stage updates of a finished request (progress stages only)"""
    return W.nodes(r) * r["stats"]["stages"]
