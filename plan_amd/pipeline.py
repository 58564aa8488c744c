"""Several snapshots through the clump-finding path at once.

The reference runs its snapshots in a loop, one block of them per MPI rank
(src/global.cpp:722-758, src/main.cpp:56-79): ReadLisFile, MakeGhostParticles,
BuildTree, FindPlanetesimals, then the catalog files.  Here the snapshots
share one GPU: each worker thread owns a device context (its own HIP stream
and device memory), so one snapshot's file reading and host-side clump logic
overlap another's kernels.  The library's calls release the GIL (ctypes), and
contexts share no state.

A series (the default) gives each snapshot the parameters the reference's
loop hands it: the box, species weights and particle masses of the first
snapshot (ReadLisFile derives them at loop_begin only,
src/tree.hpp:1563-1581), and the H an earlier snapshot's automatic choice
raised to 64 (kept in the reference's parameters, src/tree.hpp:3583-3586).
series=False treats every snapshot as the first of its own run.
"""
from __future__ import annotations

import queue
import threading
from concurrent.futures import ThreadPoolExecutor

from . import Context, PlanError


class _SeriesState:
    """What the snapshots of one series pass along: the first snapshot's
    parameters and particle count, and per snapshot whether its tree made the
    automatic choice raise H (set once its tree is built)."""

    def __init__(self, n: int):
        self.first_ready = threading.Event()
        self.first = None
        self.first_n = 0
        self.built = [threading.Event() for _ in range(n)]
        self.hop_auto = [0] * n
        self.failed = False

    def wait_first(self):
        self.first_ready.wait()
        if self.failed or self.first is None:
            raise PlanError("the series' first snapshot failed")

    def hop_before(self, k: int, hop: int) -> int:
        """H snapshot k starts from: 64 once an earlier tree raised it."""
        for j in range(k):
            self.built[j].wait()
            if self.failed:
                raise PlanError("an earlier snapshot of the series failed")
            if self.hop_auto[j]:
                return 64
        return hop


def later_params(first: dict, first_n: int, params: dict) -> dict:
    """params of a later snapshot of a series: the first snapshot's box and
    species weights, masses from its particle count."""
    p = dict(params)
    p.update(box_min=list(first["box_min"]), box_max=list(first["box_max"]),
             type_info=list(first.get("type_info", [1.0])), series_first_n=int(first_n))
    return p


def _load(ctx: Context, item, params: dict | None):
    if isinstance(item, (list, tuple)) or isinstance(item, str):
        p, dt, n = ctx.read_lis(params, item)
        return p, n
    p = item.params if params is None else params
    ctx.set_particles(p, item.pos, item.vel, item.density, item.property_index, item.id)
    return dict(p), item.n


def _run_one(ctx: Context, k: int, item, params: dict | None, write, ser: _SeriesState | None):
    base = params if params is not None or isinstance(item, (list, tuple, str)) else item.params
    if ser is not None and k > 0:
        ser.wait_first()
        base = later_params(ser.first, ser.first_n, base)
    p, n = _load(ctx, item, base)
    if ser is not None and k == 0:
        ser.first, ser.first_n = p, n
        ser.first_ready.set()
    ctx.make_ghosts()
    info = ctx.build_tree()
    if ser is not None:
        ser.hop_auto[k] = int(info.hop_auto)
        ser.built[k].set()
        hop = ser.hop_before(k, int(p["hop"]))
        ctx.set_num_hop(hop)
        p = dict(p, hop=hop)
    cat = ctx.find_planetesimals()
    if write is not None:
        write(ctx, p, cat)
    return p, cat


class SnapshotPipeline:
    """A set of device contexts kept across calls (their device memory stays
    allocated, as one context's does across snapshots)."""

    def __init__(self, contexts: int = 2, device: int = 0):
        self.contexts = [Context(device) for _ in range(max(1, contexts))]
        self._free: "queue.Queue[Context]" = queue.Queue()
        for c in self.contexts:
            self._free.put(c)

    def run(self, snapshots, params: dict | None = None, write=None, series: bool = True):
        """Clump catalogs of several snapshots, in input order.

        snapshots: each either the .lis file(s) of one snapshot (one combined
        file, or the per-process files in process order; params then holds
        the parameter-file keys) or an in-memory snapshot
        (plan_amd.synth.Snapshot).  series: the snapshots are the reference's
        loop in one process (the first one's box, weights and masses, H
        carried); False: each is a run of its own.  write(ctx, params,
        catalog), when given, runs on the worker right after
        FindPlanetesimals, while the context still holds that snapshot's tree
        (the catalog-file writers need it).  Returns [(params, catalog), ...]
        with the parameters each snapshot ran with (H included)."""
        items = list(snapshots)
        ser = _SeriesState(len(items)) if series else None

        def task(k_item):
            k, item = k_item
            ctx = self._free.get()
            try:
                return _run_one(ctx, k, item, params, write, ser)
            except BaseException:
                if ser is not None:
                    ser.failed = True
                    ser.first_ready.set()
                    for e in ser.built:
                        e.set()
                raise
            finally:
                self._free.put(ctx)

        # snapshots start in order, and one waits only for earlier ones, so
        # the workers always make progress
        with ThreadPoolExecutor(max_workers=len(self.contexts)) as ex:
            return list(ex.map(task, enumerate(items)))

    def close(self):
        for c in self.contexts:
            c.close()


def find_clumps_series(snapshots, params: dict | None = None, contexts: int = 2, device: int = 0, write=None,
                       series: bool = True):
    """SnapshotPipeline(contexts).run(snapshots, params, write, series) with
    contexts made for this call only."""
    pipe = SnapshotPipeline(contexts, device)
    try:
        return pipe.run(snapshots, params, write, series)
    finally:
        pipe.close()
