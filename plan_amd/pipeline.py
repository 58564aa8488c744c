"""Several snapshots through the clump-finding path at once.

The reference runs one snapshot per MPI rank (src/global.cpp:722-758,
src/main.cpp:56-79): ReadLisFile, MakeGhostParticles, BuildTree,
FindPlanetesimals, then the catalog files.  Here the snapshots share one
GPU: each worker thread owns a device context (its own HIP stream and
device memory), so one snapshot's file reading and host-side clump logic
overlap another's kernels.  The library's calls release the GIL (ctypes), and
contexts share no state, so the catalogs are the ones a single context gives.
"""
from __future__ import annotations

import queue
from concurrent.futures import ThreadPoolExecutor

from . import Context


def _run_one(ctx: Context, item, params: dict | None, write):
    if isinstance(item, (list, tuple)) or isinstance(item, str):
        p, dt, n = ctx.read_lis(params, item)
    else:
        p = item.params if params is None else params
        ctx.set_particles(p, item.pos, item.vel, item.density, item.property_index, item.id)
    ctx.make_ghosts()
    ctx.build_tree()
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

    def run(self, snapshots, params: dict | None = None, write=None):
        """Clump catalogs of several snapshots, in input order.

        snapshots: each either the .lis file(s) of one snapshot (one combined
        file, or the per-process files in process order; params then holds
        the parameter-file keys) or an in-memory snapshot
        (plan_amd.synth.Snapshot).  write(ctx, params, catalog), when given,
        runs on the worker right after FindPlanetesimals, while the context
        still holds that snapshot's tree (the catalog-file writers need it).
        Returns [(params, catalog), ...]."""

        def task(item):
            ctx = self._free.get()
            try:
                return _run_one(ctx, item, params, write)
            finally:
                self._free.put(ctx)

        with ThreadPoolExecutor(max_workers=len(self.contexts)) as ex:
            return list(ex.map(task, snapshots))

    def close(self):
        for c in self.contexts:
            c.close()


def find_clumps_series(snapshots, params: dict | None = None, contexts: int = 2, device: int = 0, write=None):
    """SnapshotPipeline(contexts).run(snapshots, params, write) with
    contexts made for this call only."""
    pipe = SnapshotPipeline(contexts, device)
    try:
        return pipe.run(snapshots, params, write)
    finally:
        pipe.close()
