"""ctypes binding of libamrhip.so (the C ABI in include/amr_hip.h,
include/amr_forest.h, include/amr_driver.h).

The library is the product; this module only marshals numpy arrays for the
tests and the benchmark.  Loading fails loudly when the shared library has
not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libamrhip.so")

_i = C.c_int
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_vp = C.c_void_p

# every exported entry point of the three public headers, with its signature
EXPORTS = {
    # include/amr_hip.h
    "hpc_amr_initialize_solver": (None, [_pi, _pi, _pd, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi]),
    "physics_gpu_set_params": (None, [_vp]),
    "physics_gpu_get_params": (None, [_vp]),
    "physics_go_gpu": (_i, [_pi] * 15 + [_pd, _pd] + [_pi] * 5 + [_vp, _pd, _pd, _pd]),
    "hpc_amr_evolve_solver": (_i, [_pi] * 15 + [_pd, _pd] + [_pi] * 5 + [_vp, _pd, _pd, _pd]),
    "physics_rhs_structured_gpu": (_i, [_pi, _pi, _pi, _pi, _pi, _pd, _pi, _pi, _pi, _pd, _pd]),
    "physics_time_step_gpu": (_i, [_pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd]),
    "amrgpu_create": (_vp, []),
    "amrgpu_destroy": (None, [_vp]),
    "amrgpu_last_error": (C.c_char_p, []),
    "amrgpu_setup": (_i, [_vp, _i, _i, _i, _vp, _i, _pi, _pi, _i, _pi, _pd, _pd]),
    "amrgpu_upload_soln": (_i, [_vp, _pd]),
    "amrgpu_download_soln": (_i, [_vp, _pd]),
    "amrgpu_download_quads": (_i, [_vp, _i, _i, _pd]),
    "amrgpu_rhs": (_i, [_vp, _pd]),
    "amrgpu_cfl": (_i, [_vp, _pd]),
    "amrgpu_run": (_i, [_vp, _i, _i, _pd]),
    "amrgpu_get_time": (_i, [_vp, _pd, _pd]),
    "amrgpu_set_time": (_i, [_vp, _pd]),
    "amrgpu_bench": (_i, [_vp, _i, _pd, _pd, _pi]),
    "amrgpu_num_cells": (_i, [_vp]),
    "amrgpu_num_quads": (_i, [_vp]),
    "amrgpu_copy_probe": (_i, [C.c_longlong, _i, _pd]),
    "amrgpu_set_tile": (_i, [_vp, _i, _i, _i]),
    "amrgpu_set_graphs": (_i, [_vp, _i]),
    "amrgpu_get_graphs": (_i, [_vp]),
    "amrgpu_get_tile": (_i, [_vp, _pi, _pi, _pi]),
    "amrgpu_tile_variants": (_i, [_i, _pi, _pi, _pi, _i]),
    "amrgpu_autotune": (_i, [_vp, _i, _pi, _pi, _pi, _pi, _pd]),
    "amrgpu_feature_tags": (_i, [_vp, _i, _pi, _pd, _pi]),
    "amrgpu_cfl_noalloc": (_i, [_i, _i, _i, _pd, _pd, _vp, _vp, _vp, _vp, _vp]),
    # include/amr_forest.h
    "amr_forest_new": (_vp, [_i, _i, _i, _i, _d, _d, _d, _i, _i, _i]),
    "amr_forest_destroy": (None, [_vp]),
    "amr_forest_num_quads": (_i, [_vp]),
    "amr_forest_nelem_subgrid": (_i, [_vp]),
    "amr_forest_quads": (_vp, [_vp]),
    "amr_forest_soln": (_pd, [_vp]),
    "amr_forest_tags": (_pi, [_vp]),
    "amr_forest_scale": (_d, [_vp]),
    "amr_forest_quad_geom": (None, [_vp, _i, _pd]),
    "amr_forest_build_external": (_i, [_vp, _pi, _pi, _pi, _pi, C.POINTER(_pi), C.POINTER(_pi),
                                       C.POINTER(_pd), C.POINTER(_pd)]),
    "amr_forest_refine": (_i, [_vp, _pi]),
    "amr_forest_coarsen": (_i, [_vp]),
    "amr_forest_balance": (_i, [_vp]),
    "amr_forest_is_balanced": (_i, [_vp]),
    "amr_forest_find_leaf": (_i, [_vp, _i, _i, _i]),
    "amr_forest_brick": (None, [_vp, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pi, _pi]),
    "amr_forest_balance_marked": (_i, [_vp]),
    "amr_forest_clear_marks": (None, [_vp]),
    "amr_forest_from_quads": (_vp, [_i, _i, _i, _i, _d, _d, _d, _i, _i, _i, _i, _vp]),
    "amr_forest_save": (_i, [_vp, C.c_char_p, _i, _d]),
    "amr_forest_load": (_vp, [C.c_char_p, _pi, _pd]),
    "amr_forest_write_vtk": (_i, [_vp, C.c_char_p, _i]),
    "amr_set_output_dir": (None, [C.c_char_p]),
    # include/amr_driver.h
    "amr_ic_patch": (None, [_vp, _i, _pd, _pd]),
    "amr_ic_apply": (None, [_vp, _vp]),
    "amr_ic_vortex_density": (_d, [_vp, _d, _d, _d, _d, _d, _d, _d]),
    "amr_indicator_gradient": (None, [_vp, _pd]),
    "amr_regrid": (_i, [_vp, _d, _i]),
    "amr_regrid_initial": (_i, [_vp, _vp, _d]),
    "amr_regrid_tags": (_i, [_vp, _pi, _i]),
    "amr_evolve": (_i, [_vp, _d, _i, _vp]),
    "amr_evolve_device": (_i, [_vp, _vp, _vp]),
    "amr_evolve_batch": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "amr_forest_host_parallel_enabled": (_i, []),
    "amr_forest_track_begin": (None, [_vp, _i]),
    "amr_forest_track_src": (_pi, [_vp]),
    "amr_forest_track_end": (_i, [_vp]),
    "amrgpu_indicator": (_i, [_vp, _pd]),
    "amrgpu_remesh": (_i, [_vp, _i, _pi, _pi, _i, _pi, _pi]),
    "amrgpu_remesh_nbr": (_i, [_vp, _i, _pi, _pi, _pi, _pi]),
    "amrgpu_set_geometry": (_i, [_vp, _pd]),
    "amrgpu_set_origins": (_i, [_vp, _pd]),
    "amrgpu_nbr_from_faces": (_i, [_i, _i, _pi, _pi]),
    "amr_forest_build_nbr": (_i, [_vp, _pi, _pi, _pd]),
    "amrgpu_device_state": (_i, [_vp, _vp, _vp, _pi, _pi, _vp]),
    # include/amr_tagger.h
    "amr_mlp_create": (_vp, [C.c_ulonglong]),
    "amr_mlp_destroy": (None, [_vp]),
    "amr_mlp_get_params": (_i, [_vp, _vp]),
    "amr_mlp_set_params": (_i, [_vp, _vp]),
    "amr_mlp_forward": (_i, [_vp, _vp, _i, _vp]),
    "amr_mlp_grad": (_i, [_vp, _vp, _pi, _i, _vp, _vp, _vp]),
    "amr_tagset_create": (_vp, []),
    "amr_tagset_destroy": (None, [_vp]),
    "amr_tagset_size": (_i, [_vp]),
    "amr_tagset_append": (_i, [_vp, _vp, _pi, _pd, _i]),
    "amr_tagset_get": (_i, [_vp, _i, _i, _vp, _pi, _pd]),
    "amr_tagset_append_from": (_i, [_vp, _vp, _i, _i]),
    "amrgpu_tagger_features": (_i, [_vp, _vp]),
    "amrgpu_tagger_features_host": (_i, [_vp, _vp]),
    "amr_mlp_train": (_i, [_vp, _vp, _i, _i, C.c_float, _vp, C.c_ulonglong, _vp]),
    "amr_mlp_set_refine_bias": (_i, [_vp, C.c_float]),
    "amr_mlp_predict_set": (_i, [_vp, _vp, _pi]),
    "amr_mlp_predict_ctx": (_i, [_vp, _vp, _pi]),
    "amr_tagger_collect": (_i, [_vp, _vp, _vp, _pi]),
    "amr_mlp_tagger": (_i, [_vp, _vp, _vp, _pi]),
    "amr_feature_tagger": (_i, [_vp, _vp, _vp, _pi]),
}

MLP_NPARAM = 64 * 384 + 64 + 64 * 64 + 64 + 3 * 64 + 3
TAG_COARSEN, TAG_KEEP, TAG_REFINE = 0, 1, 2


class InstParams(C.Structure):
    """amrgpu_inst_t"""
    _fields_ = [("h_base", _d), ("bc", _i * 4), ("cfl", _d), ("dt_cap", _d), ("t_end", _d),
                ("bstate", (_d * 4) * 2), ("shock_x0", _d), ("shock_angle", _d), ("shock_speed", _d)]


BC_COPY, BC_SLIP, BC_NOSLIP, BC_INFLOW, BC_SHOCK_WALL, BC_SHOCK_TOP = 0, 1, 2, 3, 4, 5


class EvolveStats(C.Structure):
    """amr_evolve_stats_t"""
    _fields_ = [("steps", _i), ("regrids", _i), ("peak_quads", _i), ("cell_stages", _d), ("t", _d),
                ("wall_gpu", _d), ("wall_regrid", _d), ("wall_tags", _d), ("wall_forest", _d),
                ("wall_remesh", _d)]


TAGGER_FN = C.CFUNCTYPE(_i, _vp, _vp, _vp, _pi)


class CollectOpts(C.Structure):
    """amr_collect_opts_t"""
    _fields_ = [("tol", _d), ("coarsen_frac", _d), ("window", _i), ("label_stride", _i),
                ("allow_coarsening", _i), ("inst", C.POINTER(InstParams))]


class MlpTaggerStruct(C.Structure):
    """amr_mlp_tagger_t"""
    _fields_ = [("mlp", _vp), ("tol", _d), ("union_plain", _i), ("predicted", C.c_longlong * 3),
                ("plain_refine", C.c_longlong), ("added_by_plain", C.c_longlong)]


MAX_FEATURE_TAGS = 8


class FeatureTaggerStruct(C.Structure):
    """amr_feature_tagger_t"""
    _fields_ = [("ntags", _i), ("methods", _i * MAX_FEATURE_TAGS), ("tols", _d * MAX_FEATURE_TAGS)]


class EvolveOpts(C.Structure):
    """amr_evolve_opts_t"""
    _fields_ = [("tol", _d), ("allow_coarsening", _i), ("regrid_interval", _i), ("tagger", _vp),
                ("tagger_ctx", _vp), ("tile_ppb", _i), ("tile_minw", _i), ("tile_design", _i),
                ("inst", C.POINTER(InstParams)), ("regrid_lag", _i), ("groups", _i)]


class PhysicsParams(C.Structure):
    """physics_gpu_params_t"""
    _fields_ = [
        ("dim", _i), ("nfields", _i), ("nelem_subgrid", _i),
        ("extern_cell_size", _i), ("extern_face_size", _i),
        ("time_scheme", _i), ("time_steps", _i), ("dt", _d),
        ("checkpoint_interval", _i), ("visualization_interval", _i), ("regrid_interval", _i),
        ("bc", _i * 6), ("cfl", _d), ("coarse_grid_h", _d), ("t_end", _d), ("verbose", _i),
    ]


class Quad(C.Structure):
    _fields_ = [("tree", _i), ("level", _i), ("x", _i), ("y", _i)]


class IC(C.Structure):
    """amr_ic_t"""
    _fields_ = [
        ("kind", _i),
        ("mach", _d), ("alpha", _d), ("beta", _d), ("gamma", _d), ("density", _d), ("pressure", _d),
        ("vortex_strength", _d), ("xo", _d), ("yo", _d),
        ("xs", _d), ("ys", _d), ("state", (_d * 4) * 4),
        ("dmr_x0", _d), ("dmr_angle", _d), ("dmr_mach", _d), ("dmr_rho", _d), ("dmr_p", _d),
    ]


IC_FREESTREAM, IC_VORTEX, IC_RIEMANN4, IC_DMR = 0, 1, 2, 3

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C amr-demo_amd` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def ipt(a):
    return _p(a, C.c_int)


def dpt(a):
    return _p(a, C.c_double)


def last_error():
    return lib().amrgpu_last_error().decode()


class GpuError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        raise GpuError(f"{what} failed rc={rc}: {last_error()}")


# ---------------------------------------------------------------------------
# host forest
# ---------------------------------------------------------------------------
class Forest:
    def __init__(self, ntx, nty, periodic, xlo, ylo, scale, n, min_level, max_level):
        self.L = lib()
        self.h = self.L.amr_forest_new(ntx, nty, int(periodic[0]), int(periodic[1]), xlo, ylo,
                                       scale, n, min_level, max_level)
        if not self.h:
            raise ValueError("amr_forest_new rejected the arguments")
        self.n = n
        self.scale = scale
        self.ntx, self.nty, self.xlo, self.ylo = ntx, nty, xlo, ylo

    @classmethod
    def _wrap(cls, h):
        f = cls.__new__(cls)
        f.L = lib()
        f.h = h
        a = [_i() for _ in range(6)]
        d = [_d() for _ in range(3)]
        f.L.amr_forest_brick(h, C.byref(a[0]), C.byref(a[1]), C.byref(a[2]), C.byref(a[3]), C.byref(d[0]),
                             C.byref(d[1]), C.byref(d[2]), C.byref(a[4]), C.byref(a[5]))
        f.ntx, f.nty, f.xlo, f.ylo, f.scale = a[0].value, a[1].value, d[0].value, d[1].value, d[2].value
        f.n = f.L.amr_forest_nelem_subgrid(h)
        return f

    @classmethod
    def load(cls, path):
        """forest from an amr_forest_save checkpoint -> (forest, istep, t)"""
        istep, t = _i(), _d()
        h = lib().amr_forest_load(os.fsencode(path), C.byref(istep), C.byref(t))
        if not h:
            raise ValueError(f"{path}: not a checkpoint of this format")
        return cls._wrap(h), istep.value, t.value

    @classmethod
    def from_quads(cls, ntx, nty, periodic, xlo, ylo, scale, n, min_level, max_level, quads):
        arr = (Quad * len(quads))(*[Quad(*q) for q in quads])
        h = lib().amr_forest_from_quads(ntx, nty, int(periodic[0]), int(periodic[1]), xlo, ylo, scale, n,
                                        min_level, max_level, len(quads), arr)
        if not h:
            raise ValueError("amr_forest_from_quads rejected the quadrants")
        return cls._wrap(h)

    def save(self, path, istep=0, t=0.0):
        if self.L.amr_forest_save(self.h, os.fsencode(path), istep, t):
            raise OSError(f"amr_forest_save({path}) failed")

    def write_vtk(self, path, subcells=False):
        if self.L.amr_forest_write_vtk(self.h, os.fsencode(path), int(subcells)):
            raise OSError(f"amr_forest_write_vtk({path}) failed")

    def __del__(self):
        try:
            if self.h:
                self.L.amr_forest_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def nquad(self):
        return self.L.amr_forest_num_quads(self.h)

    def soln(self):
        """numpy view of the forest's solution (nquad, n, n, 4) -> reference Q(4,n,n)"""
        nq, n = self.nquad, self.n
        buf = np.ctypeslib.as_array(self.L.amr_forest_soln(self.h), shape=(nq * 4 * n * n,))
        return buf.reshape(nq, n, n, 4)

    def tags(self):
        return np.ctypeslib.as_array(self.L.amr_forest_tags(self.h), shape=(self.nquad,))

    def quads(self):
        ptr = C.cast(self.L.amr_forest_quads(self.h), C.POINTER(Quad))
        return [(ptr[k].tree, ptr[k].level, ptr[k].x, ptr[k].y) for k in range(self.nquad)]

    def levels(self):
        ptr = C.cast(self.L.amr_forest_quads(self.h), C.POINTER(Quad))
        return np.array([ptr[k].level for k in range(self.nquad)], dtype=np.int32)

    def external(self):
        """(cell_info, face_info, geom, soln) copies in the reference format"""
        ncell, ng, nface, dof = _i(), _i(), _i(), _i()
        ci, fi, g, s = _pi(), _pi(), _pd(), _pd()
        rc = self.L.amr_forest_build_external(self.h, C.byref(ncell), C.byref(ng), C.byref(nface),
                                              C.byref(dof), C.byref(ci), C.byref(fi), C.byref(g),
                                              C.byref(s))
        if rc != 0:
            raise ValueError("forest is not 2:1 balanced / inconsistent")
        nc, nf, nd = ncell.value, nface.value, dof.value
        cell_info = np.ctypeslib.as_array(ci, shape=(3 * nc,)).copy()
        face_info = np.ctypeslib.as_array(fi, shape=(11 * nf,)).copy() if nf else np.zeros(0, np.int32)
        geom = np.ctypeslib.as_array(g, shape=(12 * nc,)).copy()
        soln = np.ctypeslib.as_array(s, shape=(nd,)).copy()
        return cell_info.astype(np.int32), face_info.astype(np.int32), geom, soln

    def mesh_table(self):
        """(nbr[8 nquad], level[nquad], org[2 nquad]): the device mesh table
        built straight from the forest (amr_forest_build_nbr)"""
        nq = self.nquad
        nbr = np.empty(8 * nq, np.int32)
        level = np.empty(nq, np.int32)
        org = np.empty(2 * nq, np.float64)
        rc = self.L.amr_forest_build_nbr(self.h, nbr.ctypes.data_as(_pi), level.ctypes.data_as(_pi),
                                         org.ctypes.data_as(_pd))
        if rc != 0:
            raise ValueError("forest is not 2:1 balanced / inconsistent")
        return nbr, level, org

    def set_soln(self, flat):
        v = self.soln().reshape(-1)
        v[:] = flat

    def refine(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        return self.L.amr_forest_refine(self.h, ipt(f))

    def coarsen(self):
        return self.L.amr_forest_coarsen(self.h)

    def balance(self):
        return self.L.amr_forest_balance(self.h)

    def is_balanced(self):
        return bool(self.L.amr_forest_is_balanced(self.h))

    def apply_ic(self, ic):
        self.L.amr_ic_apply(self.h, C.byref(ic))

    def indicator(self):
        ind = np.zeros(self.nquad)
        self.L.amr_indicator_gradient(self.h, dpt(ind))
        return ind

    def regrid(self, tol, allow_coarsening=True):
        return self.L.amr_regrid(self.h, tol, int(allow_coarsening))

    def regrid_initial(self, ic, tol):
        return self.L.amr_regrid_initial(self.h, C.byref(ic), tol)

    def regrid_tags(self, tags, allow_coarsening=True):
        t = np.ascontiguousarray(tags, dtype=np.int32)
        return self.L.amr_regrid_tags(self.h, ipt(t), int(allow_coarsening))

    def track_begin(self, keep_data=True):
        self.L.amr_forest_track_begin(self.h, int(keep_data))

    def track_end(self):
        src = np.ctypeslib.as_array(self.L.amr_forest_track_src(self.h), shape=(self.nquad,)).copy()
        ok = self.L.amr_forest_track_end(self.h)
        return bool(ok), src

    def evolve(self, tol, allow_coarsening=True):
        """host-regrid evolve loop (hpc_amr_evolve_solver per interval)"""
        st = EvolveStats()
        rc = self.L.amr_evolve(self.h, tol, int(allow_coarsening), C.byref(st))
        check(rc, "amr_evolve")
        return st

    def evolve_device(self, tol, regrid_interval, allow_coarsening=True, tile=None, tagger=None,
                      tagger_ctx=None, inst=None, regrid_lag=0):
        """device-resident evolve loop; tagger: optional TAGGER_FN or MlpTagger"""
        o = EvolveOpts()
        o.tol, o.allow_coarsening, o.regrid_interval = tol, int(allow_coarsening), regrid_interval
        if isinstance(tagger, (MlpTagger, FeatureTagger)):
            o.tagger = C.cast(tagger.fn(self.L), _vp)
            o.tagger_ctx = C.cast(C.pointer(tagger.s), _vp)
        else:
            o.tagger = C.cast(tagger, _vp) if tagger is not None else None
            o.tagger_ctx = tagger_ctx
        if tile == "auto":
            o.tile_ppb = -1          # amrgpu_autotune on the run's own mesh
        elif tile:
            o.tile_ppb, o.tile_minw, o.tile_design = tile
        if inst is not None:
            o.inst = C.pointer(inst)
        o.regrid_lag = regrid_lag
        st = EvolveStats()
        rc = self.L.amr_evolve_device(self.h, C.byref(o), C.byref(st))
        check(rc, "amr_evolve_device")
        return st

    def collect(self, tagset, tol, window=40, label_stride=5, coarsen_frac=0.5, allow_coarsening=True, inst=None):
        """run to the physical stop collecting labelled snapshots (amr_tagger_collect)"""
        o = CollectOpts()
        o.tol, o.coarsen_frac, o.window, o.label_stride = tol, coarsen_frac, window, label_stride
        o.allow_coarsening = int(allow_coarsening)
        if inst is not None:
            o.inst = C.pointer(inst)
        rows = C.c_int(0)
        check(self.L.amr_tagger_collect(self.h, C.byref(o), tagset.h, C.byref(rows)), "amr_tagger_collect")
        return rows.value


def evolve_batch(forests, insts, tol, regrid_interval, allow_coarsening=True, tile=None, tagger=None,
                 regrid_lag=0, groups=1):
    """amr_evolve_batch: every forest advanced in one device context
    (groups > 1: in that many contexts on their own host threads, one group's
    host regrid overlapping the others' steps); insts[k] = make_inst(...) of
    forest k.  Returns (per-instance stats, total)"""
    L = lib()
    k = len(forests)
    fa = (_vp * k)(*[f.h for f in forests])
    ia = (InstParams * k)(*insts)
    o = EvolveOpts()
    o.tol, o.allow_coarsening, o.regrid_interval = tol, int(allow_coarsening), regrid_interval
    if isinstance(tagger, (MlpTagger, FeatureTagger)):
        o.tagger = C.cast(tagger.fn(L), _vp)
        o.tagger_ctx = C.cast(C.pointer(tagger.s), _vp)
    if tile == "auto":
        o.tile_ppb = -1
    elif tile:
        o.tile_ppb, o.tile_minw, o.tile_design = tile
    o.regrid_lag = regrid_lag
    o.groups = groups
    per = (EvolveStats * k)()
    tot = EvolveStats()
    check(L.amr_evolve_batch(fa, ia, k, C.byref(o), per, C.byref(tot)), "amr_evolve_batch")
    return list(per), tot


def batch_external(forests):
    """the reference-format arrays of several forests concatenated as one
    batched mesh (forest k's quadrants after those of forests 0..k-1, face
    cell ids shifted alike, soln offsets global) and each quadrant's
    instance: (cell_info, face_info, geom, soln, cell_inst) for Gpu.setup"""
    cis, fis, gs, ss, inst = [], [], [], [], []
    off = 0
    n = forests[0].n
    for k, f in enumerate(forests):
        if f.n != n:
            raise ValueError("every forest of a batch has the same patch size")
        ci, fi, geom, soln = f.external()
        ci = ci.reshape(-1, 3).copy()
        nq = ci.shape[0]
        ci[:, 2] = 4 * n * n * (off + np.arange(nq, dtype=np.int64))
        fi = fi.reshape(-1, 11).copy()
        fi[:, 2] += off
        ids = fi[:, 3:7]
        fi[:, 3:7] = np.where(ids >= 0, ids + off, ids)
        cis.append(ci.reshape(-1)); fis.append(fi.reshape(-1)); gs.append(geom); ss.append(soln)
        inst.append(np.full(nq, k, np.int32))
        off += nq
    return (np.concatenate(cis).astype(np.int32), np.concatenate(fis).astype(np.int32), np.concatenate(gs),
            np.concatenate(ss), np.concatenate(inst))


def vortex_ic(strength=5.0, xo=5.0, yo=5.0, u=1.0, v=1.0, gamma=1.4, rho=1.0, p=1.0):
    """isentropic vortex through the reference's free-stream inputs: u = v = 1
    means |V| = sqrt(2) at 45 degrees, mach = |V| / sqrt(gamma p / rho)"""
    ic = IC()
    ic.kind = IC_VORTEX
    speed = float(np.hypot(u, v))
    ic.mach = speed / np.sqrt(gamma * p / rho)
    ic.alpha = float(np.degrees(np.arctan2(v, u)))
    ic.beta = 0.0
    ic.gamma, ic.density, ic.pressure = gamma, rho, p
    ic.vortex_strength, ic.xo, ic.yo = strength, xo, yo
    return ic


def riemann4_ic(seed, xs=0.5, ys=0.5, gamma=1.4):
    """four-quadrant Riemann states from numpy.random.default_rng(seed):
    rho log-uniform [0.1, 2], u, v uniform [-1, 1], p uniform [0.1, 2]"""
    rng = np.random.default_rng(seed)
    ic = IC()
    ic.kind = IC_RIEMANN4
    ic.gamma, ic.density, ic.pressure, ic.mach = gamma, 1.0, 1.0, 0.0
    ic.xs, ic.ys = xs, ys
    for k in range(4):
        rho = float(np.exp(rng.uniform(np.log(0.1), np.log(2.0))))
        u, v = (float(x) for x in rng.uniform(-1.0, 1.0, size=2))
        p = float(rng.uniform(0.1, 2.0))
        ic.state[k][0], ic.state[k][1], ic.state[k][2], ic.state[k][3] = rho, u, v, p
    return ic


def dmr_ic(gamma=1.4):
    """double-Mach reflection: Mach-10 shock at 60 degrees through (1/6, 0),
    pre-shock rho = 1.4, p = 1 at rest"""
    ic = IC()
    ic.kind = IC_DMR
    ic.gamma, ic.density, ic.pressure, ic.mach = gamma, 1.4, 1.0, 0.0
    ic.dmr_x0, ic.dmr_angle, ic.dmr_mach, ic.dmr_rho, ic.dmr_p = 1.0 / 6.0, 60.0, 10.0, 1.4, 1.0
    return ic


# ---------------------------------------------------------------------------
# device context
# ---------------------------------------------------------------------------
def make_inst(h_base, bc=(0, 0, 0, 0), cfl=0.3, dt_cap=1.0, t_end=0.0, bstate=None, shock=None):
    ip = InstParams()
    ip.h_base = h_base
    for s in range(4):
        ip.bc[s] = int(bc[s])
    ip.cfl, ip.dt_cap, ip.t_end = cfl, dt_cap, t_end
    if bstate is not None:
        for a in range(2):
            for f in range(4):
                ip.bstate[a][f] = float(bstate[a][f])
    if shock is not None:
        ip.shock_x0, ip.shock_angle, ip.shock_speed = (float(v) for v in shock)
    return ip


def dmr_states(gamma=1.4, mach=10.0, rho1=1.4, p1=1.0, angle=60.0):
    """post- and pre-shock conservative states of the double-Mach reflection
    (Rankine-Hugoniot for a shock of Mach `mach` into gas at rest, moving
    along the normal of a line at `angle` degrees to the x axis), its speed"""
    th = np.radians(angle)
    c1 = np.sqrt(gamma * p1 / rho1)
    w = mach * c1
    rho2 = rho1 * (gamma + 1.0) * mach * mach / ((gamma - 1.0) * mach * mach + 2.0)
    p2 = p1 * (2.0 * gamma * mach * mach - (gamma - 1.0)) / (gamma + 1.0)
    un = w * (1.0 - rho1 / rho2)
    u2, v2 = un * np.sin(th), -un * np.cos(th)
    post = (rho2, rho2 * u2, rho2 * v2, p2 / (gamma - 1.0) + 0.5 * rho2 * (u2 * u2 + v2 * v2))
    pre = (rho1, 0.0, 0.0, p1 / (gamma - 1.0))
    return post, pre, w


def dmr_inst(h_base=1.0, cfl=0.3, dt_cap=1.0, t_end=0.0, x0=1.0 / 6.0, angle=60.0):
    """walls of the double-Mach reflection: post-shock inflow on the left,
    outflow on the right, post-shock then reflecting wall (from x0) along the
    bottom, the moving shock along the top"""
    post, pre, w = dmr_states(angle=angle)
    return make_inst(h_base, (BC_INFLOW, BC_COPY, BC_SHOCK_WALL, BC_SHOCK_TOP), cfl, dt_cap, t_end,
                     bstate=(post, pre), shock=(x0, angle, w))


class Gpu:
    """amrgpu_ctx"""

    def __init__(self):
        self.L = lib()
        self.h = self.L.amrgpu_create()
        if not self.h:
            raise GpuError(f"amrgpu_create failed: {last_error()}")
        self.ninst = 0

    def close(self):
        if self.h:
            self.L.amrgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self, n, scheme, insts, cell_info, face_info, geom, soln, cell_inst=None):
        ninst = len(insts)
        arr = (InstParams * ninst)(*insts)
        ci = np.ascontiguousarray(cell_info, dtype=np.int32)
        fi = np.ascontiguousarray(face_info, dtype=np.int32)
        g = np.ascontiguousarray(geom, dtype=np.float64)
        s = np.ascontiguousarray(soln, dtype=np.float64)
        cinst = None if cell_inst is None else np.ascontiguousarray(cell_inst, dtype=np.int32)
        ncell = ci.size // 3
        nface = fi.size // 11
        rc = self.L.amrgpu_setup(self.h, n, scheme, ninst, C.cast(arr, _vp), ncell, ipt(ci),
                                 ipt(cinst) if cinst is not None else None, nface,
                                 ipt(fi) if nface else ipt(np.zeros(11, np.int32)), dpt(g), dpt(s))
        check(rc, "amrgpu_setup")
        self.ninst = ninst
        self.ndof = s.size
        self.n2 = n * n

    def indicator(self):
        out = np.empty(self.ndof // 4 // self.n2)
        check(self.L.amrgpu_indicator(self.h, dpt(out)), "amrgpu_indicator")
        return out

    def feature_tags(self, methods, tols):
        """amrgpu_feature_tags: the reference's density-threshold tags
        (method 1 below, 2 above) per quadrant of the current state"""
        m = np.ascontiguousarray(methods, dtype=np.int32)
        t = np.ascontiguousarray(tols, dtype=np.float64)
        out = np.empty(self.ndof // 4 // self.n2, np.int32)
        check(self.L.amrgpu_feature_tags(self.h, m.size, ipt(m), dpt(t), ipt(out)), "amrgpu_feature_tags")
        return out

    def remesh(self, cell_info, face_info, src, cell_inst=None):
        ci = np.ascontiguousarray(cell_info, dtype=np.int32)
        fi = np.ascontiguousarray(face_info, dtype=np.int32)
        sr = np.ascontiguousarray(src, dtype=np.int32)
        cinst = None if cell_inst is None else np.ascontiguousarray(cell_inst, dtype=np.int32)
        rc = self.L.amrgpu_remesh(self.h, ci.size // 3, ipt(ci), ipt(cinst) if cinst is not None else None,
                                  fi.size // 11, ipt(fi), ipt(sr))
        check(rc, "amrgpu_remesh")
        self.ndof = ci.size // 3 * 4 * self.n2

    def remesh_nbr(self, level, nbr, src, cell_inst=None):
        """amrgpu_remesh_nbr: the new mesh as the device table of
        Forest.mesh_table (levels, nbr[8 nquad]) instead of cell_info / face_info"""
        lv = np.ascontiguousarray(level, dtype=np.int32)
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        sr = np.ascontiguousarray(src, dtype=np.int32)
        cinst = None if cell_inst is None else np.ascontiguousarray(cell_inst, dtype=np.int32)
        rc = self.L.amrgpu_remesh_nbr(self.h, lv.size, ipt(lv), ipt(cinst) if cinst is not None else None,
                                      ipt(nb), ipt(sr))
        check(rc, "amrgpu_remesh_nbr")
        self.ndof = lv.size * 4 * self.n2

    def set_origins(self, org):
        check(self.L.amrgpu_set_origins(self.h, dpt(np.ascontiguousarray(org, dtype=np.float64))),
              "amrgpu_set_origins")

    def rhs(self):
        out = np.empty(self.ndof)
        check(self.L.amrgpu_rhs(self.h, dpt(out)), "amrgpu_rhs")
        return out

    def cfl(self):
        out = np.empty(self.ninst)
        check(self.L.amrgpu_cfl(self.h, dpt(out)), "amrgpu_cfl")
        return out

    def run(self, nsteps, graph=True):
        rn = _d()
        rc = self.L.amrgpu_run(self.h, nsteps, int(graph), C.byref(rn))
        if rc == 3:
            raise FloatingPointError("nan detected in residual")
        check(rc, "amrgpu_run")
        return rn.value

    def soln(self):
        out = np.empty(self.ndof)
        check(self.L.amrgpu_download_soln(self.h, dpt(out)), "amrgpu_download_soln")
        return out

    def upload(self, soln):
        s = np.ascontiguousarray(soln, dtype=np.float64)
        check(self.L.amrgpu_upload_soln(self.h, dpt(s)), "amrgpu_upload_soln")

    def time(self):
        t = np.empty(self.ninst)
        dt = np.empty(self.ninst)
        check(self.L.amrgpu_get_time(self.h, dpt(t), dpt(dt)), "amrgpu_get_time")
        return t, dt

    def set_time(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        check(self.L.amrgpu_set_time(self.h, dpt(t)), "amrgpu_set_time")

    def bench(self, nsteps):
        ms = _d()
        st = (C.c_double * 8)()
        ns = _i()
        check(self.L.amrgpu_bench(self.h, nsteps, C.byref(ms), st, C.byref(ns)), "amrgpu_bench")
        return ms.value, [st[k] for k in range(ns.value)]

    def set_tile(self, ppb, minw=1, design=0):
        check(self.L.amrgpu_set_tile(self.h, ppb, minw, design), "amrgpu_set_tile")

    def set_graphs(self, enable):
        """step launch mode: True replays a captured HIP graph per step,
        False launches the stages directly (the default)"""
        check(self.L.amrgpu_set_graphs(self.h, 1 if enable else 0), "amrgpu_set_graphs")

    def graphs(self):
        return bool(self.L.amrgpu_get_graphs(self.h))

    def tile(self):
        p, m, d = _i(), _i(), _i()
        check(self.L.amrgpu_get_tile(self.h, C.byref(p), C.byref(m), C.byref(d)), "amrgpu_get_tile")
        return p.value, m.value, d.value

    def autotune(self, nsteps=20):
        nv = _i()
        ppb, mw, ds, ms = (_i * 32)(), (_i * 32)(), (_i * 32)(), (_d * 32)()
        check(self.L.amrgpu_autotune(self.h, nsteps, C.byref(nv), ppb, mw, ds, ms), "amrgpu_autotune")
        return [(ppb[k], mw[k], ds[k], ms[k]) for k in range(nv.value)]

    @property
    def ncells(self):
        return self.L.amrgpu_num_cells(self.h)

    def nquad(self):
        return self.L.amrgpu_num_quads(self.h)

    def tagger_features(self):
        X = np.zeros((self.nquad(), 384), np.uint16)
        check(self.L.amrgpu_tagger_features_host(self.h, X.ctypes.data), "amrgpu_tagger_features_host")
        return X


def set_output_dir(path):
    """amr_set_output_dir: where the evolve loops write VTK and checkpoints (None: nowhere)"""
    lib().amr_set_output_dir(os.fsencode(path) if path else None)


def set_params(**kw):
    p = PhysicsParams()
    lib().physics_gpu_get_params(C.byref(p))
    for k, v in kw.items():
        if k == "bc":
            for s in range(6):
                p.bc[s] = int(v[s]) if s < len(v) else 1
        else:
            setattr(p, k, v)
    lib().physics_gpu_set_params(C.byref(p))
    return p


def get_params():
    p = PhysicsParams()
    lib().physics_gpu_get_params(C.byref(p))
    return p


def nbr_from_faces(ncell, face_info):
    """the device mesh table the library derives from reference face_info"""
    fi = np.ascontiguousarray(face_info, dtype=np.int32)
    nbr = np.empty(8 * ncell, np.int32)
    check(lib().amrgpu_nbr_from_faces(ncell, len(fi) // 11, fi.ctypes.data_as(_pi), nbr.ctypes.data_as(_pi)),
          "amrgpu_nbr_from_faces")
    return nbr


def tile_variants(n):
    a, b, c = (_i * 32)(), (_i * 32)(), (_i * 32)()
    k = lib().amrgpu_tile_variants(n, a, b, c, 32)
    return [(a[i], b[i], c[i]) for i in range(min(k, 32))]


def copy_probe(nbytes, reps=20):
    g = _d()
    check(lib().amrgpu_copy_probe(int(nbytes), reps, C.byref(g)), "amrgpu_copy_probe")
    return g.value


# ---------------------------------------------------------------------------
# learned tagger
# ---------------------------------------------------------------------------
class Mlp:
    """amr_mlp_t: 384-64-64-3 BF16 perceptron on the MFMA units"""

    def __init__(self, seed=0):
        self.L = lib()
        self.h = self.L.amr_mlp_create(seed)
        if not self.h:
            raise GpuError(f"amr_mlp_create failed: {last_error()}")

    def close(self):
        if self.h:
            self.L.amr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        P = np.zeros(MLP_NPARAM, np.float32)
        check(self.L.amr_mlp_get_params(self.h, P.ctypes.data), "amr_mlp_get_params")
        return P

    def set_params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        assert P.size == MLP_NPARAM
        check(self.L.amr_mlp_set_params(self.h, P.ctypes.data), "amr_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.uint16)
        out = np.zeros((X.shape[0], 3), np.float32)
        check(self.L.amr_mlp_forward(self.h, X.ctypes.data, X.shape[0], out.ctypes.data), "amr_mlp_forward")
        return out

    def grad(self, X, y, cw=(1.0, 1.0, 1.0)):
        X = np.ascontiguousarray(X, np.uint16)
        y = np.ascontiguousarray(y, np.int32)
        cw = np.ascontiguousarray(cw, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_float(0.0)
        check(self.L.amr_mlp_grad(self.h, X.ctypes.data, ipt(y), X.shape[0], cw.ctypes.data, g.ctypes.data,
                                  C.byref(loss)), "amr_mlp_grad")
        return loss.value, g

    def train(self, tagset, epochs, batch=4096, lr=1e-3, cw=None, seed=0):
        loss = np.zeros(max(epochs, 1), np.float32)
        cwp = None
        if cw is not None:
            cwa = np.ascontiguousarray(cw, np.float32)
            cwp = cwa.ctypes.data
        check(self.L.amr_mlp_train(self.h, tagset.h, epochs, batch, lr, cwp, seed, loss.ctypes.data),
              "amr_mlp_train")
        return loss[:epochs]

    def set_refine_bias(self, bias):
        check(self.L.amr_mlp_set_refine_bias(self.h, float(bias)), "amr_mlp_set_refine_bias")

    def predict(self, tagset):
        out = np.zeros(tagset.size, np.int32)
        check(self.L.amr_mlp_predict_set(self.h, tagset.h, ipt(out)), "amr_mlp_predict_set")
        return out

    def predict_ctx(self, gpu):
        out = np.zeros(gpu.nquad(), np.int32)
        check(self.L.amr_mlp_predict_ctx(self.h, gpu.h, ipt(out)), "amr_mlp_predict_ctx")
        return out


class Tagset:
    """amr_tagset_t: labelled feature rows in device memory"""

    def __init__(self):
        self.L = lib()
        self.h = self.L.amr_tagset_create()

    def close(self):
        if self.h:
            self.L.amr_tagset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self):
        return self.L.amr_tagset_size(self.h)

    def append(self, X, y, ind=None):
        X = np.ascontiguousarray(X, np.uint16)
        y = np.ascontiguousarray(y, np.int32)
        indp = dpt(np.ascontiguousarray(ind, float)) if ind is not None else None
        check(self.L.amr_tagset_append(self.h, X.ctypes.data, ipt(y), indp, X.shape[0]), "amr_tagset_append")

    def append_from(self, other, first=0, n=None):
        n = other.size - first if n is None else n
        check(self.L.amr_tagset_append_from(self.h, other.h, first, n), "amr_tagset_append_from")

    def get(self, first=0, n=None, features=True):
        n = self.size - first if n is None else n
        X = np.zeros((n, 384), np.uint16) if features else None
        y = np.zeros(n, np.int32)
        ind = np.zeros(n)
        check(self.L.amr_tagset_get(self.h, first, n, X.ctypes.data if features else None, ipt(y), dpt(ind)),
              "amr_tagset_get")
        return X, y, ind


class MlpTagger:
    """amr_mlp_tagger_t for Forest.evolve_device(tagger=...)"""

    def __init__(self, mlp, tol, union_plain=True):
        self.mlp = mlp
        self.s = MlpTaggerStruct()
        self.s.mlp = mlp.h
        self.s.tol = tol
        self.s.union_plain = int(union_plain)

    @staticmethod
    def fn(L):
        return L.amr_mlp_tagger

    def stats(self):
        return {"predicted": list(self.s.predicted), "plain_refine": self.s.plain_refine,
                "added_by_plain": self.s.added_by_plain}


class FeatureTagger:
    """amr_feature_tagger_t for evolve_device / evolve_batch(tagger=...): the
    reference's density-threshold tags, methods[k] in {1: below, 2: above}"""

    def __init__(self, methods, tols):
        if len(methods) != len(tols) or len(methods) > MAX_FEATURE_TAGS:
            raise ValueError("methods and tols: equal lengths, at most 8")
        self.s = FeatureTaggerStruct()
        self.s.ntags = len(methods)
        for k, (m, t) in enumerate(zip(methods, tols)):
            self.s.methods[k] = int(m)
            self.s.tols[k] = float(t)

    @staticmethod
    def fn(L):
        return L.amr_feature_tagger
