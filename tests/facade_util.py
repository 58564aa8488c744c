"""Builds tests/cpp/facade_run (the reference-shaped C++ host program over
the C ABI) and writes its snapshot files."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_run.cpp")


def build_facade_run(out_dir):
    exe = os.path.join(out_dir, "facade_run")
    lib_dir = os.path.join(ROOT, "plan_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
                    "-L", lib_dir, "-lplan_hip", f"-Wl,-rpath,{lib_dir}", "-o", exe], check=True)
    return exe


def write_snapshot(path, snap):
    from plan_amd.abi import make_params

    p = make_params(snap.params)
    with open(path, "wb") as f:
        f.write(np.uint32(snap.n).tobytes())
        f.write(bytes(p))
        f.write(np.ascontiguousarray(snap.pos, np.float64).tobytes())
        f.write(np.ascontiguousarray(snap.vel, np.float64).tobytes())
        f.write(np.ascontiguousarray(snap.density, np.float64).tobytes())
        f.write(np.ascontiguousarray(snap.property_index, np.int32).tobytes())
        f.write(np.ascontiguousarray(snap.id, np.uint32).tobytes())


def read_catalog(path):
    with open(path) as f:
        head = f.readline().split()
        n, mp_tot = int(head[1]), float(head[3])
        clumps = []
        for _ in range(n):
            peak, nm, mass, hill = f.readline().split()
            ids = np.array(f.readline().split(), dtype=np.uint32)
            clumps.append(dict(peak_index=int(peak), num_members=int(nm), total_mass=float(mass),
                               Hill_radius=float(hill), ids=ids))
    return dict(Mp_tot=mp_tot, clumps=clumps)
