"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
hot-path build (oracle/_ref/libamrref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference that GPU results are compared
with (and, for the baseline, timed).  The product never loads it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
ORACLE_NATIVE_SO = os.path.join(HERE, "liboracle_native.so")
REF_SO = os.path.join(HERE, "_ref", "libamrref.so")

_i, _d = C.c_int, C.c_double
_pi, _pd, _vp = C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_void_p


def ipt(a):
    return a.ctypes.data_as(_pi)


def dpt(a):
    return a.ctypes.data_as(_pd)


class OrcParams(C.Structure):
    _fields_ = [("nelem_subgrid", _i), ("time_scheme", _i), ("bc", _i * 6), ("cfl", _d),
                ("dt", _d), ("coarse_grid_h", _d), ("bstate", (_d * 4) * 2), ("shock_x0", _d),
                ("shock_angle", _d), ("shock_speed", _d), ("time", _d)]


RHS_FN = C.CFUNCTYPE(None, _vp, _pd, _pd)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = C.CDLL(path)
        sig = {
            "orc_flux_roe_2d": [_pd, _pd, _pd, _pd],
            "orc_flux_bc_2d": [_pd, _pd, _pd],
            "orc_bc_2d": [_i, _pd, _pd, _pd],
            "orc_rhs_structured": [_vp, _i, _pi, _pd, _i, _pi, _i, _pd, _pd],
            "orc_time_step": [_vp, _i, _pi, _pd, _pd, _pd],
            "orc_run_steps": [_vp, _i, _pi, _pd, _i, _pi, _i, _i, _pd, _pd],
            "orc_run_steps_threaded": [_vp, _i, _pi, _pd, _i, _pi, _i, _i, _pd, _pd, _i],
            "orc_rk_step": [_vp, _i, _pi, _pd, _i, _pi, _i, _d, _pd, _pd],
            "orc_initial_condition_2d": [_i, _i, _pd, _pd] + [_d] * 9,
            "orc_refine_2d": [_i, _pd, _pd, _pd, _pd, _pd],
            "orc_coarsen_2d": [_i, _pd, _pd, _pd, _pd, _pd],
            "orc_tag_feature": [_i, _pi, _pd, _i, _pd],
            "orc_norm_L2_residual": [_i, _i, _pi, _pd],
            "orc_rk_generic": [_i, _i, _d, _pd, _pd, RHS_FN, _vp],
            "orc_bf16": [C.c_float],
            "orc_tagger_features": [_i, _i, _i, _pi, _pd, _vp],
            "orc_mlp_forward": [_vp, _vp, _i, _i, _pd],
            "orc_mlp_grad": [_vp, _vp, _pi, _i, _vp, _i, _pd],
        }
        res = {"orc_run_steps": _d, "orc_run_steps_threaded": _d, "orc_rk_step": _d,
               "orc_tag_feature": _i, "orc_norm_L2_residual": _d, "orc_rk_generic": _d, "orc_bf16": C.c_ushort,
               "orc_tagger_features": _i, "orc_mlp_grad": _d}
        for k, a in sig.items():
            f = getattr(L, k)
            f.argtypes = a
            f.restype = res.get(k)
        self.L = L

    @staticmethod
    def params(n, bc=(0, 0, 0, 0), cfl=0.3, dt=1.0, h_base=1.0, scheme=3, bstate=None, shock=None, time=0.0):
        """bstate: two conservative states, shock: (x0, angle_deg, speed) for
        the project wall types 3-5; time: instance time of the residual"""
        p = OrcParams()
        p.nelem_subgrid, p.time_scheme = n, scheme
        for s in range(6):
            p.bc[s] = int(bc[s]) if s < len(bc) else 1
        p.cfl, p.dt, p.coarse_grid_h = cfl, dt, h_base
        if bstate is not None:
            for a in range(2):
                for f in range(4):
                    p.bstate[a][f] = float(bstate[a][f])
        if shock is not None:
            p.shock_x0, p.shock_angle, p.shock_speed = (float(v) for v in shock)
        p.time = time
        return p

    def flux(self, ql, qr, nrm):
        out = np.zeros(4)
        self.L.orc_flux_roe_2d(dpt(np.ascontiguousarray(ql, float)), dpt(np.ascontiguousarray(qr, float)),
                               dpt(np.ascontiguousarray(nrm, float)), dpt(out))
        return out

    def flux_bc(self, q, nrm):
        out = np.zeros(4)
        self.L.orc_flux_bc_2d(dpt(np.ascontiguousarray(q, float)), dpt(np.ascontiguousarray(nrm, float)), dpt(out))
        return out

    def bc(self, t, nrm, q):
        out = np.zeros(4)
        self.L.orc_bc_2d(t, dpt(np.ascontiguousarray(nrm, float)), dpt(np.ascontiguousarray(q, float)), dpt(out))
        return out

    def rhs(self, p, cell_info, face_info, geom, soln):
        ci = np.ascontiguousarray(cell_info, np.int32)
        fi = np.ascontiguousarray(face_info, np.int32)
        g = np.ascontiguousarray(geom, float)
        s = np.ascontiguousarray(soln, float)
        out = np.zeros_like(s)
        self.L.orc_rhs_structured(C.byref(p), ci.size // 3, ipt(ci), dpt(g), fi.size // 11, ipt(fi),
                                  s.size, dpt(s), dpt(out))
        return out

    def norm(self, n, cell_info, res):
        """sum of squares of res over the real patches of cell_info (no sqrt)"""
        ci = np.ascontiguousarray(cell_info, np.int32)
        r = np.ascontiguousarray(res, float)
        return self.L.orc_norm_L2_residual(n, ci.size // 3, ipt(ci), dpt(r))

    def time_step(self, p, cell_info, geom, soln, dt0):
        ci = np.ascontiguousarray(cell_info, np.int32)
        g = np.ascontiguousarray(geom, float)
        s = np.ascontiguousarray(soln, float)
        dt = C.c_double(dt0)
        self.L.orc_time_step(C.byref(p), ci.size // 3, ipt(ci), dpt(g), dpt(s), C.byref(dt))
        return dt.value

    def run(self, p, cell_info, face_info, geom, soln, nsteps, threads=0):
        ci = np.ascontiguousarray(cell_info, np.int32)
        fi = np.ascontiguousarray(face_info, np.int32)
        g = np.ascontiguousarray(geom, float)
        q = np.array(soln, dtype=float, copy=True)
        dt = C.c_double(0.0)
        if threads:
            rn = self.L.orc_run_steps_threaded(C.byref(p), ci.size // 3, ipt(ci), dpt(g), fi.size // 11,
                                               ipt(fi), q.size, nsteps, dpt(q), C.byref(dt), threads)
        else:
            rn = self.L.orc_run_steps(C.byref(p), ci.size // 3, ipt(ci), dpt(g), fi.size // 11, ipt(fi),
                                      q.size, nsteps, dpt(q), C.byref(dt))
        return q, rn, dt.value

    def rk_generic(self, scheme, dt, q0, rhs, with_norm=False):
        """one step of scheme on dq/dt = rhs(q) (rhs: numpy -> numpy); with
        with_norm also the step's residual norm (sqrt of the last stage's sum
        of squares)"""
        q = np.array(q0, dtype=float, copy=True)
        work = np.zeros(5 * q.size)
        n = q.size

        def cb(_ctx, qp, kp):
            qa = np.ctypeslib.as_array(qp, shape=(n,))
            ka = np.ctypeslib.as_array(kp, shape=(n,))
            ka[:] = rhs(qa.copy())

        f = RHS_FN(cb)
        rn = self.L.orc_rk_generic(scheme, n, dt, dpt(q), dpt(work), f, None)
        return (q, rn) if with_norm else q

    def ic(self, kind, n, geom, mach, alpha, beta, gamma, density, pressure, beta_l, xo, yo):
        q = np.zeros(4 * n * n)
        self.L.orc_initial_condition_2d(kind, n, dpt(np.ascontiguousarray(geom, float)), dpt(q), mach,
                                        alpha, beta, gamma, density, pressure, beta_l, xo, yo)
        return q

    # ---- learned tagger (no reference counterpart) ----
    def tagger_features(self, n, cell_info, face_info, soln):
        ci = np.ascontiguousarray(cell_info, np.int32)
        fi = np.ascontiguousarray(face_info, np.int32)
        s = np.ascontiguousarray(soln, float)
        ncell = ci.size // 3
        X = np.zeros((ncell, 384), np.uint16)
        rc = self.L.orc_tagger_features(n, ncell, fi.size // 11, ipt(fi), dpt(s), X.ctypes.data)
        if rc:
            raise ValueError("orc_tagger_features failed")
        return X

    def mlp_forward(self, P, X, exact=False):
        P = np.ascontiguousarray(P, np.float32)
        X = np.ascontiguousarray(X, np.uint16)
        out = np.zeros((X.shape[0], 3))
        self.L.orc_mlp_forward(P.ctypes.data, X.ctypes.data, X.shape[0], int(exact), dpt(out))
        return out

    def mlp_grad(self, P, X, y, cw, exact=False):
        P = np.ascontiguousarray(P, np.float32)
        X = np.ascontiguousarray(X, np.uint16)
        y = np.ascontiguousarray(y, np.int32)
        cw = np.ascontiguousarray(cw, np.float32)
        g = np.zeros(P.size)
        loss = self.L.orc_mlp_grad(P.ctypes.data, X.ctypes.data, ipt(y), X.shape[0], cw.ctypes.data,
                                   int(exact), dpt(g))
        return loss, g


# ---------------------------------------------------------------------------
# the reference's own build
# ---------------------------------------------------------------------------
class PhysicsInterface(C.Structure):
    """physics_interface_t (src/physics/src/physics_var_interface.h:20-37)"""
    _fields_ = [("dim", _i), ("nfringe", _i), ("nfields", _i), ("nelem_subgrid", _i),
                ("extern_cell_size", _i), ("extern_face_size", _i), ("extern_edge_size", _i),
                ("extern_node_size", _i), ("extern_mpi_comm_size", _i), ("nlevels", _i),
                ("coarse_grid_h", _d)]


class PhysicsFluid(C.Structure):
    """physics_fluid_t (src/physics/src/physics_var_physics.h:21-47)"""
    _fields_ = [("initial_condition", _i), ("bc", _i * 6), ("amr_ntag_methods", _i),
                ("amr_tag_methods", _pi), ("amr_tag_tolerances", _pd), ("cfl", _d), ("mach", _d),
                ("alpha", _d), ("beta", _d), ("gamma", _d), ("density", _d), ("pressure", _d)]


def ref_available():
    return os.path.exists(REF_SO)


class Reference:
    """The reference's Fortran kernels + C glue, compiled from its sources."""

    def __init__(self):
        if not ref_available():
            raise RuntimeError("oracle/_ref not built (make -C oracle ref, needs the reference sources)")
        L = C.CDLL(REF_SO)
        self.L = L
        self.iface = PhysicsInterface.in_dll(L, "d_physics_interface")
        self.fluid = PhysicsFluid.in_dll(L, "d_physics_fluid")
        self._flux = getattr(L, "_QMflux_2d_modulePflux_roe_2d")
        self._flux.argtypes = [_pd] * 4
        self._fluxbc = getattr(L, "_QMflux_2d_modulePflux_bc_2d")
        self._fluxbc.argtypes = [_pd] * 3
        self._bc = getattr(L, "_QMbc_2d_modulePbc_2d")
        self._bc.argtypes = [_pi, _pd, _pd, _pd]
        self._ic = getattr(L, "_QMinitial_condition_2d_modulePinitial_condition_2d")
        self._ic.argtypes = [_pi, _pi, _pi, _pi, _pd, _pd, _pd] + [_pd] * 6
        L.physics_rhs_structured.argtypes = [_pi, _pi, _pi, _pi, _pi, _pd, _pi, _pi, _pi, _pd, _pd]
        L.physics_time_step.argtypes = [_pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd]
        L.physics_norm_L2_residual.argtypes = [_i, _i, _i, _i, _pi, _pd]
        L.physics_norm_L2_residual.restype = _d
        L.solver_tag_feature.argtypes = [_pi, _pi, _pd, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pi]
        L.solver_refine_operator_2d.argtypes = [_pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd, _pd, _pd, _pd]
        L.solver_coarsen_operator_2d.argtypes = [_pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd, _pd, _pd, _pd]

    def configure(self, n, bc=(0, 0, 0, 0), cfl=0.3, h_base=1.0):
        self.iface.dim, self.iface.nfields, self.iface.nelem_subgrid = 2, 4, n
        self.iface.extern_cell_size, self.iface.extern_face_size = 3, 11
        self.iface.coarse_grid_h = h_base
        for s in range(6):
            self.fluid.bc[s] = int(bc[s]) if s < len(bc) else 1
        self.fluid.cfl = cfl

    def flux(self, ql, qr, nrm):
        out = np.zeros(4)
        self._flux(dpt(np.ascontiguousarray(ql, float)), dpt(np.ascontiguousarray(qr, float)),
                   dpt(np.ascontiguousarray(nrm, float)), dpt(out))
        return out

    def flux_bc(self, q, nrm):
        out = np.zeros(4)
        self._fluxbc(dpt(np.ascontiguousarray(q, float)), dpt(np.ascontiguousarray(nrm, float)), dpt(out))
        return out

    def bc(self, t, nrm, q):
        out = np.zeros(4)
        tt = C.c_int(t)
        self._bc(C.byref(tt), dpt(np.ascontiguousarray(nrm, float)), dpt(np.ascontiguousarray(q, float)),
                 dpt(out))
        return out

    def rhs(self, n, cell_info, face_info, geom, soln):
        ci = np.ascontiguousarray(cell_info, np.int32)
        fi = np.ascontiguousarray(face_info, np.int32)
        g = np.ascontiguousarray(geom, float)
        s = np.ascontiguousarray(soln, float)
        out = np.zeros_like(s)
        dim, nf, nn = C.c_int(2), C.c_int(4), C.c_int(n)
        ncell, nface, dof = C.c_int(ci.size // 3), C.c_int(fi.size // 11), C.c_int(s.size)
        self.L.physics_rhs_structured(C.byref(dim), C.byref(nf), C.byref(nn), C.byref(ncell), ipt(ci),
                                      dpt(g), C.byref(nface), ipt(fi), C.byref(dof), dpt(s), dpt(out))
        return out

    def time_step(self, n, cell_info, geom, soln, dt0):
        ci = np.ascontiguousarray(cell_info, np.int32)
        g = np.ascontiguousarray(geom, float)
        s = np.ascontiguousarray(soln, float)
        dim, nf, nn, ncell = C.c_int(2), C.c_int(4), C.c_int(n), C.c_int(ci.size // 3)
        dt = C.c_double(dt0)
        self.L.physics_time_step(C.byref(dim), C.byref(nf), C.byref(nn), C.byref(ncell), ipt(ci), dpt(g),
                                 dpt(s), C.byref(dt))
        return dt.value

    def ic(self, kind, n, h, geom, mach, alpha, beta, gamma, density, pressure):
        q = np.zeros(4 * n * n)
        a = [C.c_int(kind), C.c_int(0), C.c_int(4), C.c_int(n)]
        d = [C.c_double(x) for x in (h, mach, alpha, beta, gamma, density, pressure)]
        self._ic(C.byref(a[0]), C.byref(a[1]), C.byref(a[2]), C.byref(a[3]), C.byref(d[0]),
                 dpt(np.ascontiguousarray(geom, float)), dpt(q), *[C.byref(x) for x in d[1:]])
        return q

    def norm(self, n, cell_info, res):
        """physics_norm_L2_residual: sum of squares (no sqrt) over the real
        patches named by cell_info"""
        ci = np.ascontiguousarray(cell_info, np.int32)
        r = np.ascontiguousarray(res, float)
        return self.L.physics_norm_L2_residual(2, 4, n, ci.size // 3, ipt(ci), dpt(r))

    def tag(self, methods, tols, n, q):
        m = np.ascontiguousarray(methods, np.int32)
        t = np.ascontiguousarray(tols, float)
        nt, dim, un, nf, nn, ct = (C.c_int(len(m)), C.c_int(2), C.c_int(0), C.c_int(4), C.c_int(n), C.c_int(0))
        h = C.c_double(1.0)
        xyz = np.zeros(12)
        tag = C.c_int(-1)
        self.L.solver_tag_feature(C.byref(nt), ipt(m), dpt(t), C.byref(dim), C.byref(un), C.byref(nf),
                                  C.byref(nn), C.byref(ct), C.byref(h), dpt(xyz),
                                  dpt(np.ascontiguousarray(q, float)), C.byref(tag))
        return tag.value

    def refine(self, n, parent):
        kids = [np.zeros(4 * n * n) for _ in range(4)]
        un, nf, nn, ct = C.c_int(0), C.c_int(4), C.c_int(n), C.c_int(0)
        h = C.c_double(1.0)
        xyz = np.zeros(12)
        self.L.solver_refine_operator_2d(C.byref(un), C.byref(nf), C.byref(nn), C.byref(ct), C.byref(h),
                                         dpt(xyz), dpt(np.ascontiguousarray(parent, float)),
                                         *[dpt(k) for k in kids])
        return kids

    def coarsen(self, n, c1, c2, c3, c4):
        out = np.zeros(4 * n * n)
        un, nf, nn, ct = C.c_int(0), C.c_int(4), C.c_int(n), C.c_int(0)
        h = C.c_double(1.0)
        xyz = np.zeros(12)
        self.L.solver_coarsen_operator_2d(C.byref(un), C.byref(nf), C.byref(nn), C.byref(ct), C.byref(h),
                                          dpt(xyz), *[dpt(np.ascontiguousarray(c, float)) for c in
                                                      (c1, c2, c3, c4)], dpt(out))
        return out
