"""ctypes mirrors of the POD types in include/plan_params.h."""
from __future__ import annotations

import ctypes as C

PLAN_MAX_TYPES = 16


class PlanParams(C.Structure):
    _fields_ = [
        ("box_min", C.c_double * 3),
        ("box_max", C.c_double * 3),
        ("time", C.c_double),
        ("solid_to_gas_ratio", C.c_double),
        ("four_pi_G", C.c_double),
        ("clump_diffuse_threshold", C.c_double),
        ("Hill_fraction_for_merge", C.c_double),
        ("type_info", C.c_double * PLAN_MAX_TYPES),
        ("box_resolution", C.c_uint32 * 3),
        ("num_types", C.c_uint32),
        ("num_knn", C.c_uint32),
        ("num_hop", C.c_uint32),
        ("fixed_num_hop", C.c_int32),
        ("num_peaks", C.c_uint32),
        ("no_ghost", C.c_int32),
        ("series_first_n", C.c_uint32),
    ]


class PlanDerived(C.Structure):
    _fields_ = [
        ("box_center", C.c_double * 3),
        ("box_length", C.c_double * 3),
        ("box_half_width", C.c_double * 3),
        ("max_half_width", C.c_double),
        ("ghost_zone_width", C.c_double * 3),
        ("max_ghost_zone_width", C.c_double),
        ("cell_length", C.c_double * 3),
        ("cell_volume", C.c_double),
        ("mass_total_code_units", C.c_double),
        ("grav_constant", C.c_double),
        ("shear_speed", C.c_double),
        ("q", C.c_double),
        ("Omega", C.c_double),
        ("min_trusted_mass_code_unit", C.c_double),
        ("mass_per_particle", C.c_double * PLAN_MAX_TYPES),
    ]


class PlanClump(C.Structure):
    _fields_ = [
        ("peak_index", C.c_uint32),
        ("num_members", C.c_uint32),
        ("num_particles", C.c_uint32),
        ("member_offset", C.c_uint32),
        ("total_mass", C.c_double),
        ("center_of_mass", C.c_double * 3),
        ("vel_com", C.c_double * 3),
        ("Hill_radius", C.c_double),
        ("J", C.c_double * 3),
        ("J_quarter_Hill", C.c_double * 3),
        ("one10th_radius", C.c_double),
        ("two_sigma_mass_radius", C.c_double),
        ("max_distance", C.c_double),
        ("peak_density", C.c_double),
    ]


class PlanCatalog(C.Structure):
    _fields_ = [
        ("num_clumps", C.c_uint32),
        ("num_members_total", C.c_uint32),
        ("clumps", C.POINTER(PlanClump)),
        ("members", C.POINTER(C.c_uint32)),
        ("member_original_ids", C.POINTER(C.c_uint32)),
        ("density_threshold", C.c_double),
        ("peak_density_threshold", C.c_double),
        ("clump_mass_threshold", C.c_double),
        ("Mp_max", C.c_double),
        ("Mp_tot", C.c_double),
    ]


def make_params(d: dict) -> PlanParams:
    """Fill a PlanParams from the dict produced by plan_amd.synth.default_params."""
    p = PlanParams()
    for k in range(3):
        p.box_min[k] = d["box_min"][k]
        p.box_max[k] = d["box_max"][k]
        p.box_resolution[k] = int(d["Nx"][k])
    p.time = float(d.get("time", 0.0))
    p.solid_to_gas_ratio = float(d["Z"])
    p.four_pi_G = float(d["four_pi_G"])
    p.clump_diffuse_threshold = float(d.get("clump_diffuse_threshold", -1.0))
    p.Hill_fraction_for_merge = float(d.get("Hill_fraction_for_merge", -1.0))
    ti = d.get("type_info", [1.0])
    p.num_types = len(ti)
    for i, v in enumerate(ti):
        p.type_info[i] = float(v)
    p.num_knn = int(d["knn"])
    p.num_hop = int(d["hop"])
    p.fixed_num_hop = int(d.get("fixed_hop", 0))
    p.num_peaks = int(d.get("num_peaks", 0))
    p.no_ghost = int(d.get("no_ghost", 0))
    p.series_first_n = int(d.get("series_first_n", 0))
    return p


def catalog_to_py(cat: PlanCatalog) -> dict:
    """Copy a catalog into plain Python/numpy containers."""
    import numpy as np

    n = cat.num_clumps
    tot = cat.num_members_total
    clumps = []
    for i in range(n):
        c = cat.clumps[i]
        clumps.append(dict(
            peak_index=c.peak_index, num_members=c.num_members, num_particles=c.num_particles,
            member_offset=c.member_offset, total_mass=c.total_mass,
            center_of_mass=tuple(c.center_of_mass), vel_com=tuple(c.vel_com),
            Hill_radius=c.Hill_radius, J=tuple(c.J), J_quarter_Hill=tuple(c.J_quarter_Hill),
            one10th_radius=c.one10th_radius, two_sigma_mass_radius=c.two_sigma_mass_radius,
            max_distance=c.max_distance, peak_density=c.peak_density))
    members = np.ctypeslib.as_array(cat.members, shape=(max(tot, 1),))[:tot].copy()
    oids = np.ctypeslib.as_array(cat.member_original_ids, shape=(max(tot, 1),))[:tot].copy()
    return dict(clumps=clumps, members=members, member_original_ids=oids,
                density_threshold=cat.density_threshold,
                peak_density_threshold=cat.peak_density_threshold,
                clump_mass_threshold=cat.clump_mass_threshold,
                Mp_max=cat.Mp_max, Mp_tot=cat.Mp_tot)


def py_to_catalog(d: dict):
    """A PlanCatalog over a dict from catalog_to_py; returns (catalog, keep),
    keep holding the arrays the catalog points into."""
    import numpy as np

    n = len(d["clumps"])
    arr = (PlanClump * max(n, 1))()
    for i, c in enumerate(d["clumps"]):
        o = arr[i]
        for k in ("peak_index", "num_members", "num_particles", "member_offset", "total_mass", "Hill_radius",
                  "one10th_radius", "two_sigma_mass_radius", "max_distance", "peak_density"):
            setattr(o, k, c[k])
        for k in ("center_of_mass", "vel_com", "J", "J_quarter_Hill"):
            for j in range(3):
                getattr(o, k)[j] = c[k][j]
    mem = np.ascontiguousarray(d["members"], np.uint32)
    oid = np.ascontiguousarray(d["member_original_ids"], np.uint32)
    cat = PlanCatalog()
    cat.num_clumps = n
    cat.num_members_total = len(mem)
    cat.clumps = C.cast(arr, C.POINTER(PlanClump))
    cat.members = mem.ctypes.data_as(C.POINTER(C.c_uint32))
    cat.member_original_ids = oid.ctypes.data_as(C.POINTER(C.c_uint32))
    for k in ("density_threshold", "peak_density_threshold", "clump_mass_threshold", "Mp_max", "Mp_tot"):
        setattr(cat, k, d[k])
    return cat, (arr, mem, oid)


class PlanBoundnessIn(C.Structure):
    _fields_ = [("num_clumps", C.c_uint32), ("offsets", C.c_void_p), ("pos", C.c_void_p), ("vel", C.c_void_p),
                ("mass", C.c_void_p), ("G", C.c_double), ("Omega", C.c_double), ("q", C.c_double)]


class PlanBoundnessClump(C.Structure):
    _fields_ = [("n_in", C.c_uint32), ("n_bound", C.c_uint32), ("iterations", C.c_uint32), ("reserved", C.c_uint32),
                ("mass", C.c_double), ("center_of_mass", C.c_double * 3), ("vel_com", C.c_double * 3),
                ("E_kin", C.c_double), ("E_pot", C.c_double)]


class PlanBoundnessStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("pair_evaluations", C.c_double), ("refit_clumps", C.c_uint32),
                ("graph_replays", C.c_uint32), ("tile_shape", C.c_int32), ("reserved", C.c_int32),
                ("tune_ms", C.c_double * 4)]
