"""Continuous-batching serving loop over the native server (chq_server_*): arriving runs are
admitted against an HBM budget and per-batch slots, join the batch of their scheme between
steps, and leave it when they reach tlim / nlim. `play()` drives a timed arrival stream
(e.g. seeded Poisson) through it and reports time-to-solution percentiles and runs/s.
"""
import ctypes as C
import time

from ._capi import check, lib
from .mesh import Mesh

STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "released"}
PRECISION = {None: -1, "deck": -1, "fp64": 0, "mixed": 1}


class Server:
    def __init__(self, hbm_budget_bytes, slots=64, chunk=4):
        L = lib()
        h = C.c_void_p()
        check(L.chq_server_create(C.c_size_t(int(hbm_budget_bytes)), int(slots), int(chunk), C.byref(h)),
              "chq_server_create")
        self.h, self.L = h, L

    def close(self):
        if getattr(self, "h", None):
            self.L.chq_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, deck_text, overrides=(), precision=None):
        jid = C.c_int()
        check(self.L.chq_server_submit(self.h, deck_text.encode(), "\n".join(overrides).encode(),
                                       PRECISION[precision], C.byref(jid)), "chq_server_submit")
        return jid.value

    def pump(self, max_done=4096):
        ids = (C.c_int * max_done)()
        n = C.c_int()
        check(self.L.chq_server_pump(self.h, max_done, ids, C.byref(n)), "chq_server_pump")
        return [ids[i] for i in range(min(n.value, max_done))]

    def job(self, jid):
        info = (C.c_double * 10)()
        check(self.L.chq_server_job(self.h, jid, info), "chq_server_job")
        return {"state": STATES[int(info[0])], "precision": "mixed" if info[1] else "fp64",
                "fell_back": bool(info[2]), "cycle": int(info[3]), "t": info[4], "t_submit": info[5],
                "t_admit": info[6], "t_done": info[7], "bytes": int(info[8]), "fofc_total": int(info[9]),
                "error": self.L.chq_server_job_error(self.h, jid).decode()}

    def mesh(self, jid):
        """Non-owning view of a finished (or running) run's context, to download its state."""
        h = self.L.chq_server_job_ctx(self.h, jid)
        if not h:
            raise KeyError(f"job {jid} has no device state")
        return Mesh.borrow(h)

    def release(self, jid):
        check(self.L.chq_server_release(self.h, jid), "chq_server_release")

    def stats(self):
        s = (C.c_double * 11)()
        check(self.L.chq_server_stats(self.h, s), "chq_server_stats")
        return {"queued": int(s[0]), "running": int(s[1]), "finished": int(s[2]), "bytes_in_use": int(s[3]),
                "budget": int(s[4]), "batches": int(s[5]), "batch_steps": int(s[6]), "host_launch_s": s[7],
                "host_admit_s": s[8], "host_poll_wait_s": s[9], "host_retire_s": s[10]}

    def clock(self):
        return self.L.chq_server_clock()

    def play(self, arrivals, on_done=None, release=True, idle_sleep=1e-3):
        """Serve `arrivals` = [(arrival offset in s, deck text, overrides, precision)] in real
        time: each run is submitted once its offset has passed, the loop pumps while anything is
        queued or running. on_done(job_id, server) runs before a finished run is released.
        Returns per-run records (arrival offset, time-to-solution, queue wait, ...)."""
        order = sorted(range(len(arrivals)), key=lambda i: arrivals[i][0])
        t0 = time.perf_counter()
        nxt = 0
        jobs = {}
        recs = {}
        pending = 0
        while nxt < len(order) or pending:
            now = time.perf_counter() - t0
            while nxt < len(order) and arrivals[order[nxt]][0] <= now:
                i = order[nxt]
                _, text, ov, prec = arrivals[i]
                jid = self.submit(text, ov, prec)
                jobs[jid] = (i, time.perf_counter() - t0)
                pending += 1
                nxt += 1
            if not pending:
                time.sleep(min(idle_sleep, max(0.0, arrivals[order[nxt]][0] - now)))
                continue
            for jid in self.pump():
                i, t_arr = jobs[jid]
                info = self.job(jid)
                t_fin = time.perf_counter() - t0
                recs[i] = {"job": jid, "arrival": t_arr, "finish": t_fin, "tts": t_fin - t_arr,
                           "wait": info["t_admit"] - info["t_submit"], "cycle": info["cycle"],
                           "precision": info["precision"], "fell_back": info["fell_back"],
                           "state": info["state"], "bytes": info["bytes"]}
                if on_done is not None:
                    on_done(jid, self)
                if release:
                    self.release(jid)
                pending -= 1
        return [recs[i] for i in range(len(arrivals))]
