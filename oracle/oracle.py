"""TEST INFRASTRUCTURE ONLY: ctypes front end of the FP64 CPU oracle.

The oracle restates the reference algorithm (see sbslpm_oracle.cpp for the
file:line of every function).  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import this module; the product path never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sbslpm.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_ppd = ctypes.POINTER(_dp)


def build():
    subprocess.check_call([os.path.join(_HERE, "build.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for f in ("orc_panel_count", "orc_leaf_count", "orc_interaction_count"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int
        L.orc_get_panels.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_get_interactions.argtypes = [ctypes.c_void_p, _ip]
        L.orc_interp_init.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_cc_weights_textbook.argtypes = [ctypes.c_int, _dp]
        L.orc_interp_vals_bli.argtypes = [_dp] + [ctypes.c_double] * 6 + [ctypes.c_int]
        L.orc_interp_vals_bli.restype = ctypes.c_int
        L.orc_face_from_xyz.argtypes = [ctypes.c_double] * 3
        L.orc_face_from_xyz.restype = ctypes.c_int
        L.orc_xyz_from_xieta.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_xyz_from_xieta.restype = ctypes.c_int
        L.orc_xieta_from_xyz.argtypes = [ctypes.c_double] * 3 + [_dp]
        L.orc_xieta_from_xyz_face.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _dp]
        L.orc_xieta_from_xyz_face.restype = ctypes.c_int
        L.orc_sphere_tri_area.argtypes = [_dp, _dp, _dp]
        L.orc_sphere_tri_area.restype = ctypes.c_double
        L.orc_point_locate_panel.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
        L.orc_point_locate_panel.restype = ctypes.c_int
        L.orc_spec_points.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_proxy_points.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_upward_pass.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_upward_pass.restype = ctypes.c_int
        L.orc_downward_pass.argtypes = [ctypes.c_void_p, ctypes.c_int, _ppd, _ppd, _ppd]
        L.orc_downward_pass.restype = ctypes.c_int
        L.orc_bve_vel_interactions.argtypes = [ctypes.c_void_p, _ppd, _ppd, _dp, _dp, ctypes.c_double, _ppd, _ppd]
        L.orc_bve_vel_interactions_sample.argtypes = [ctypes.c_void_p, _ppd, _ppd, _dp, _dp, ctypes.c_double,
                                                      _ppd, _ppd, ctypes.c_int]
        L.orc_poisson_fmm_interactions.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_direct_sum_inv_lap.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_int64]
        L.orc_direct_sum_bve_vel.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _ppd, ctypes.c_double, _ppd,
                                             ctypes.c_int64, ctypes.c_int64]
        L.orc_direct_sum_inv_lap_targets.argtypes = [ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int64, _dp, _dp, _dp,
                                                     _dp, _dp, _dp]
        L.orc_direct_sum_bve_vel_targets.argtypes = [ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int64, _dp, _dp, _dp,
                                                     _dp, _dp, ctypes.c_double, _ppd]
        L.orc_bve_displacement.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _ppd, _dp, ctypes.c_double,
                                           ctypes.c_double, _dp, _ppd, _ppd]
        L.orc_bve_displacement.restype = ctypes.c_int
        L.orc_bve_find_departure_points.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _ppd, ctypes.c_double,
                                                    ctypes.c_int, _dp, _dp, _dp]
        L.orc_bve_find_departure_points.restype = ctypes.c_int
        L.orc_bve_departure_to_target.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                                  ctypes.c_double, _dp, _dp]
        L.orc_bve_departure_to_target.restype = ctypes.c_int
        L.orc_bve_back_rk4_step.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ppd]
        L.orc_bve_back_rk4_step.restype = ctypes.c_int
        L.orc_swe_vel_interactions.argtypes = [ctypes.c_void_p, ctypes.c_int, _ppd, _ppd, _dp, _dp, _dp, _dp,
                                               ctypes.c_double, _ppd, _ppd]
        L.orc_swe_departure_to_target.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _ppd, _dp, ctypes.c_int, _ppd,
                                                  _dp]
        L.orc_swe_departure_to_target.restype = ctypes.c_int
        L.orc_panel_laplacian.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_velocity_gradient.argtypes = [ctypes.c_void_p, _ppd, ctypes.c_int, _ppd]
        L.orc_swe_tendency.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _ppd, _dp, _dp, _dp, _dp, ctypes.c_double,
                                       ctypes.c_int, _dp, _dp, _dp]
        L.orc_apply_topography.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _ppd, _ppd, _dp, ctypes.c_int,
                                           ctypes.c_int, _dp]
        L.orc_laplacian_interactions.argtypes = [ctypes.c_void_p, _ppd, _ppd, _dp, _dp, _dp, _dp]
        L.orc_laplacian_interactions.restype = None
        L.orc_swe_height_correction.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _ppd, _dp, _dp, ctypes.c_double,
                                                _dp]
        L.orc_swe_back_rk4_step.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, _ppd, _dp]
        L.orc_eval_targets.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int64,
                                       _dp, _dp, _dp, _ppd]
        L.orc_eval_targets.restype = ctypes.c_int
        L.orc_interp_points.argtypes = [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int, _ppd, _ppd]
        L.orc_interp_points.restype = ctypes.c_int
        L.orc_xyz_vel_to_uv_vel.argtypes = [ctypes.c_int64] + [_dp] * 8
        L.orc_dedupe.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _ip, _ip, _dp, _dp, _dp]
        L.orc_dedupe.restype = ctypes.c_int64
        L.orc_unify_boundary_vals.argtypes = [ctypes.c_int64, _ip, _ip, ctypes.c_int64, _dp]
        L.orc_unify_boundary_vals.restype = ctypes.c_int
        L.orc_irreg_create.restype = ctypes.c_void_p
        L.orc_irreg_create.argtypes = [ctypes.c_int64, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_irreg_destroy.argtypes = [ctypes.c_void_p]
        L.orc_irreg_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        L.orc_irreg_panels.argtypes = [ctypes.c_void_p, _ip, _dp, _ip]
        L.orc_irreg_interactions.argtypes = [ctypes.c_void_p, _ip]
        L.orc_irreg_upward_pass.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_irreg_upward_pass.restype = ctypes.c_int
        L.orc_irreg_downward_pass_3.argtypes = [ctypes.c_void_p, _ppd, _ppd]
        L.orc_irreg_downward_pass_3.restype = ctypes.c_int
        L.orc_irreg_swe_vel_interactions.argtypes = [ctypes.c_void_p, _ppd, _ppd, _dp, _dp, _dp, _dp, _dp,
                                                     ctypes.c_double]
        L.orc_set_swe_physics.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 5 + [ctypes.c_int]
        L.orc_set_swe_physics.restype = None
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _i, _d = ctypes.c_int, ctypes.c_double
        L.orc_ll_tendencies.argtypes = [_i, _i, _d, _d] + [_dp] * 13
        L.orc_ll_tendencies.restype = ctypes.c_int
        L.orc_ll_polar_filter.argtypes = [_i, _i, _d] + [_dp] * 4
        L.orc_ll_arrival_values.argtypes = [_i, _i, _d, _d] + [_dp] * 14
        L.orc_ll_arrival_values.restype = ctypes.c_int
        L.orc_ll_field_averaging.argtypes = [_i, _i] + [_dp] * 5 + [_d]
        L.orc_ll_laplacian_diffusion.argtypes = [_i, _i, _d, _d] + [_dp] * 5
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _pp(arrs):
    if arrs is None:
        return None
    ptrs = (_dp * len(arrs))(*[_p(a) for a in arrs])
    _pp.keep.append((ptrs, arrs))
    if len(_pp.keep) > 64:
        del _pp.keep[:32]
    return ctypes.cast(ptrs, _ppd)


_pp.keep = []


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def interp_init(degree, weight_rule=0):
    tab = np.zeros(((degree + 1) ** 2, 4))
    lib().orc_interp_init(degree, weight_rule, _p(tab))
    return tab


def cc_weights_textbook(degree):
    w = np.zeros(degree + 1)
    lib().orc_cc_weights_textbook(degree, _p(w))
    return w


def interp_vals_bli(xi, eta, min_xi, max_xi, min_eta, max_eta, degree):
    basis = np.zeros((degree + 1) ** 2)
    err = lib().orc_interp_vals_bli(_p(basis), xi, eta, min_xi, max_xi, min_eta, max_eta, degree)
    return basis, err


def face_from_xyz(x, y, z):
    return lib().orc_face_from_xyz(x, y, z)


def xyz_from_xieta(xi, eta, face):
    out = np.zeros(3)
    err = lib().orc_xyz_from_xieta(xi, eta, face, _p(out))
    return out, err


def xieta_from_xyz(x, y, z, face=None):
    out = np.zeros(2)
    if face is None:
        lib().orc_xieta_from_xyz(x, y, z, _p(out))
        return out
    err = lib().orc_xieta_from_xyz_face(x, y, z, face, _p(out))
    return out, err


def sphere_tri_area(a, b, c):
    return lib().orc_sphere_tri_area(_p(np.asarray(a, float)), _p(np.asarray(b, float)), _p(np.asarray(c, float)))


def direct_sum_inv_lap(x, y, z, area, pots, lb=0, ub=None):
    n = len(x)
    ub = n if ub is None else ub
    soln = np.zeros(n)
    lib().orc_direct_sum_inv_lap(n, _p(x), _p(y), _p(z), _p(area), _p(pots), _p(soln), lb, ub)
    return soln


def direct_sum_bve_vel(x, y, z, area, vor, eps, disp=None, lb=0, ub=None):
    n = len(x)
    ub = n if ub is None else ub
    vel = [np.zeros(n) for _ in range(3)]
    lib().orc_direct_sum_bve_vel(n, _p(x), _p(y), _p(z), _p(area), _p(vor), _pp(disp), eps, _pp(vel), lb, ub)
    return vel


def direct_sum_inv_lap_targets(tx, ty, tz, sx, sy, sz, area, pots):
    out = np.zeros(len(tx))
    lib().orc_direct_sum_inv_lap_targets(len(tx), _p(tx), _p(ty), _p(tz), len(sx), _p(sx), _p(sy), _p(sz),
                                         _p(area), _p(pots), _p(out))
    return out


def direct_sum_bve_vel_targets(tx, ty, tz, sx, sy, sz, area, vor, eps):
    """Biot-Savart velocity at separate targets over undisplaced sources (kernel of
    bve_vel_impl.hpp:63-66); returns [vx, vy, vz]."""
    vel = [np.zeros(len(tx)) for _ in range(3)]
    lib().orc_direct_sum_bve_vel_targets(len(tx), _p(tx), _p(ty), _p(tz), len(sx), _p(sx), _p(sy), _p(sz),
                                         _p(area), _p(vor), float(eps), _pp(vel))
    return vel


class OracleFMM:
    """Tree + interaction list + passes of the reference FMM on the host.

    levels L (leaves at level L-1), proxy degree d, particle-grid degree
    md = m-1 (md == d is the reference layout), MAC theta, particle weight rule
    (0 reference quadrature, 1 Clenshaw-Curtis, 2 equal area).
    """

    def __init__(self, levels, degree, leaf_degree=None, theta=0.7, weight_rule=0):
        self.levels, self.d = levels, degree
        self.md = degree if leaf_degree is None else leaf_degree
        self.theta, self.weight_rule = theta, weight_rule
        self.h = lib().orc_create(levels, degree, self.md, theta, weight_rule)
        if not self.h:
            raise ValueError("oracle: bad parameters")
        self.P = lib().orc_panel_count(self.h)
        self.Pl = lib().orc_leaf_count(self.h)
        self.offset = self.P - self.Pl
        self.n = (degree + 1) ** 2
        self.m2 = (self.md + 1) ** 2

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def interaction_count(self):
        return lib().orc_interaction_count(self.h)

    def panels(self):
        ints = np.zeros((self.P, 8), dtype=np.int32)
        dbls = np.zeros((self.P, 6))
        lib().orc_get_panels(self.h, ints.ctypes.data_as(_ip), _p(dbls))
        return ints, dbls

    def interactions(self):
        out = np.zeros((self.interaction_count, 3), dtype=np.int32)
        lib().orc_get_interactions(self.h, out.ctypes.data_as(_ip))
        return out

    def spec_points(self):
        a = [np.zeros((self.Pl, self.m2)) for _ in range(4)]
        lib().orc_spec_points(self.h, *[_p(v) for v in a])
        return a

    def proxy_points(self):
        a = [np.zeros((self.P, self.n)) for _ in range(3)]
        lib().orc_proxy_points(self.h, *[_p(v) for v in a])
        return a

    def point_locate(self, x, y, z):
        return lib().orc_point_locate_panel(self.h, x, y, z)

    def upward_pass(self, area, pots):
        qp = np.zeros((self.Pl, self.m2))
        qx = np.zeros((self.P, self.n))
        err = lib().orc_upward_pass(self.h, _p(area), _p(pots), _p(qp), _p(qx))
        if err:
            raise RuntimeError("oracle upward_pass: geometry error")
        return qp, qx

    def downward_pass(self, phi_proxy, phi_part=None):
        ncomp = len(phi_proxy)
        sols = [np.zeros((self.Pl, self.m2)) for _ in range(ncomp)]
        err = lib().orc_downward_pass(self.h, ncomp, _pp(phi_proxy), _pp(phi_part), _pp(sols))
        if err:
            raise RuntimeError("oracle downward_pass: geometry error")
        return sols

    def bve_vel_interactions(self, q_part, q_proxy, eps, disp_part=None, disp_proxy=None):
        if disp_part is None:
            disp_part = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        if disp_proxy is None:
            disp_proxy = [np.zeros((self.P, self.n)) for _ in range(3)]
        pp = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        px = [np.zeros((self.P, self.n)) for _ in range(3)]
        lib().orc_bve_vel_interactions(self.h, _pp(disp_part), _pp(disp_proxy), _p(q_part), _p(q_proxy), eps,
                                       _pp(pp), _pp(px))
        return pp, px

    def bve_vel_interactions_sample(self, q_part, q_proxy, eps, stride):
        """Interactions of target panels t % stride == 0 only (timing samples)."""
        dp = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        dx = [np.zeros((self.P, self.n)) for _ in range(3)]
        pp = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        px = [np.zeros((self.P, self.n)) for _ in range(3)]
        lib().orc_bve_vel_interactions_sample(self.h, _pp(dp), _pp(dx), _p(q_part), _p(q_proxy), eps, _pp(pp),
                                              _pp(px), int(stride))
        return pp, px

    def poisson_fmm_interactions(self, q_part, q_proxy):
        pp = np.zeros((self.Pl, self.m2))
        px = np.zeros((self.P, self.n))
        lib().orc_poisson_fmm_interactions(self.h, _p(q_part), _p(q_proxy), _p(pp), _p(px))
        return pp, px

    def bve_velocity(self, area, vor, eps, disp_part=None, disp_proxy=None):
        qp, qx = self.upward_pass(area, vor)
        pp, px = self.bve_vel_interactions(qp, qx, eps, disp_part, disp_proxy)
        return self.downward_pass(px, pp)

    def poisson_solve(self, area, f):
        qp, qx = self.upward_pass(area, f)
        pp, px = self.poisson_fmm_interactions(qp, qx)
        return self.downward_pass([px], [pp])[0]

    def swe_vel_interactions(self, variant, qz_part, qz_proxy, qd_part, qd_proxy, eps, disp_part=None,
                             disp_proxy=None):
        """swe_vel_interactions (variant 1) / swe_vel_interactions_2 (variant 2); returns (phi_part, phi_proxy)."""
        if disp_part is None:
            disp_part = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        if disp_proxy is None:
            disp_proxy = [np.zeros((self.P, self.n)) for _ in range(3)]
        pp = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        px = [np.zeros((self.P, self.n)) for _ in range(3)]
        lib().orc_swe_vel_interactions(self.h, variant, _pp(disp_part), _pp(disp_proxy), _p(qz_part), _p(qz_proxy),
                                       _p(qd_part), _p(qd_proxy), eps, _pp(pp), _pp(px))
        return pp, px

    def swe_velocity(self, variant, area, vor, div, eps, disp_part=None, disp_proxy=None):
        """Upward passes of vor and div, SWE interactions, downward pass (swe_time_step.cpp:655-672)."""
        qzp, qzx = self.upward_pass(area, vor)
        qdp, qdx = self.upward_pass(area, div)
        pp, px = self.swe_vel_interactions(variant, qzp, qzx, qdp, qdx, eps, disp_part, disp_proxy)
        return self.downward_pass(px, pp)

    def eval_targets(self, kind, area, f, tx, ty, tz, eps=0.0):
        """FMM value at target points (kind "poisson": psi, "bve": velocity
        components); returns a list of 1 or 3 arrays."""
        k = {"poisson": 0, "bve": 1}[kind]
        tx, ty, tz = [np.ascontiguousarray(v, dtype=np.float64) for v in (tx, ty, tz)]
        out = [np.zeros(tx.size) for _ in range(1 if k == 0 else 3)]
        err = lib().orc_eval_targets(self.h, k, _p(area), _p(f), eps, tx.size, _p(tx), _p(ty), _p(tz), _pp(out))
        if err:
            raise RuntimeError("oracle eval_targets: geometry error")
        return out

    def bve_displacement(self, x, y, z, vel, vor, tau, omega):
        nv = np.zeros((self.Pl, self.m2))
        dp = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        dx = [np.zeros((self.P, self.n)) for _ in range(3)]
        err = lib().orc_bve_displacement(self.h, _p(x), _p(y), _p(z), _pp(vel), _p(vor), tau, omega, _p(nv),
                                         _pp(dp), _pp(dx))
        if err:
            raise RuntimeError("oracle bve_displacement: geometry error")
        return nv, dp, dx

    def bve_find_departure_points(self, x, y, z, vels, dt, combine=0):
        dep = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        flat = [v for stage in vels for v in stage]
        err = lib().orc_bve_find_departure_points(self.h, _p(x), _p(y), _p(z), _pp(flat), dt, combine,
                                                  *[_p(d) for d in dep])
        if err:
            raise RuntimeError("oracle departure points: geometry error")
        return dep

    def bve_departure_to_target(self, z, dep, vor, omega, tracers=None):
        ntr = 0 if tracers is None else tracers.shape[-1]
        tr = np.zeros((self.Pl, self.m2, max(ntr, 1))) if tracers is None else tracers
        nv = np.zeros((self.Pl, self.m2))
        ntrac = np.zeros_like(tr)
        err = lib().orc_bve_departure_to_target(self.h, _p(z), *[_p(d) for d in dep], _p(vor), _p(tr), ntr,
                                                omega, _p(nv), _p(ntrac))
        if err:
            raise RuntimeError("oracle departure_to_target: geometry error")
        return nv, (ntrac if ntr else None)

    def swe_departure_to_target(self, dep, fields, tracers=None):
        """swe_departure_to_target: fields = (vor, div, h); returns (3 new fields, new tracers)."""
        ntr = 0 if tracers is None else tracers.shape[-1]
        tr = np.zeros((self.Pl, self.m2, max(ntr, 1))) if tracers is None else tracers
        out = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        ntrac = np.zeros_like(tr)
        err = lib().orc_swe_departure_to_target(self.h, *[_p(d) for d in dep], _pp(list(fields)), _p(tr), ntr,
                                                _pp(out), _p(ntrac))
        if err:
            raise RuntimeError("oracle swe_departure_to_target: geometry error")
        return out, (ntrac if ntr else None)

    def bve_back_rk4_step(self, x, y, z, area, vor, dt, omega, eps, combine=0, tracers=None):
        """Returns (new vorticity, stage-0 velocity, new tracers)."""
        vor = vor.copy()
        ntr = 0 if tracers is None else tracers.shape[-1]
        tr = np.zeros((self.Pl, self.m2, 1)) if tracers is None else tracers.copy()
        vel = [np.zeros((self.Pl, self.m2)) for _ in range(3)]
        err = lib().orc_bve_back_rk4_step(self.h, _p(x), _p(y), _p(z), _p(area), _p(vor), _p(tr), ntr, dt, omega,
                                          eps, combine, _pp(vel))
        if err:
            raise RuntimeError("oracle bve_back_rk4_step: geometry error")
        return vor, vel, (tr if ntr else None)

    # ---- shallow-water stepper (swe_time_step.cpp) ----
    def _f(self, a):
        return np.ascontiguousarray(a, dtype=np.float64).reshape(self.Pl, self.m2)

    def _zp(self, k=None):
        return np.zeros((self.Pl, self.m2)) if k is None else [np.zeros((self.Pl, self.m2)) for _ in range(k)]

    def panel_laplacian(self, f):
        """laplacian (deriv_funcs.cpp:27-37) on every leaf's particle grid."""
        lap = self._zp()
        lib().orc_panel_laplacian(self.h, _p(self._f(f)), _p(lap))
        return lap

    def velocity_gradient(self, vel, rule=0):
        """[9] leaf gradients d vel_a / d x_b at 3a+b (single_panel_grad, deriv_funcs_impl.hpp:272-305)."""
        g = self._zp(9)
        if lib().orc_velocity_gradient(self.h, _pp([self._f(v) for v in vel]), rule, _pp(g)):
            raise RuntimeError("oracle velocity_gradient: geometry error")
        return g

    def set_swe_physics(self, gravity=1.0, mountain=(0.0, 1.0, 0.0, 0.0), form=0):
        """Coefficient of the free-surface Laplacian in the divergence tendency (1 in the
        reference; g/a^2 for heights in metres), the Gaussian mountain of topography 2:
        (height, width in radians, centre latitude, centre longitude), and the step form
        bits (1: regularised stage-1 Laplacian, 2: free-surface height correction)."""
        lib().orc_set_swe_physics(self.h, float(gravity), *[float(v) for v in mountain], int(form))

    def swe_tendency(self, x, y, z, vel, vor, div, height, lap, omega, rule=0):
        """swe_tendency_computation (swe_time_step.cpp:76-161): (vor_t, div_t, h_t)."""
        out = self._zp(3)
        err = lib().orc_swe_tendency(self.h, _p(x), _p(y), _p(z), _pp([self._f(v) for v in vel]), _p(self._f(vor)),
                                     _p(self._f(div)), _p(self._f(height)), _p(self._f(lap)), omega, rule,
                                     *[_p(o) for o in out])
        if err:
            raise RuntimeError("oracle swe_tendency: geometry error")
        return out

    def apply_topography(self, x, y, z, height, topo=1, rule=0, disp_part=None, disp_proxy=None):
        """apply_topography / cone_mountain (topography.cpp:6-56)."""
        dp = self._zp(3) if disp_part is None else [self._f(d) for d in disp_part]
        dx = [np.zeros((self.P, self.n)) for _ in range(3)] if disp_proxy is None else \
            [np.ascontiguousarray(d, dtype=np.float64) for d in disp_proxy]
        eff = self._zp()
        if lib().orc_apply_topography(self.h, _p(x), _p(y), _p(z), _pp(dp), _pp(dx), _p(self._f(height)), topo,
                                      rule, _p(eff)):
            raise RuntimeError("oracle apply_topography: bad argument")
        return eff

    def laplacian_interactions(self, q_part, q_proxy, disp_part=None, disp_proxy=None):
        """laplacian_interactions (laplacian.cpp:6-35); returns (phi_part, phi_proxy)."""
        dp = self._zp(3) if disp_part is None else [self._f(d) for d in disp_part]
        dx = [np.zeros((self.P, self.n)) for _ in range(3)] if disp_proxy is None else \
            [np.ascontiguousarray(d, dtype=np.float64) for d in disp_proxy]
        pp, px = self._zp(), np.zeros((self.P, self.n))
        lib().orc_laplacian_interactions(self.h, _pp(dp), _pp(dx), _p(self._f(q_part)),
                                         _p(np.ascontiguousarray(q_proxy, dtype=np.float64)), _p(pp), _p(px))
        return pp, px

    def swe_height_correction(self, x, y, z, vel, height, h_t, dt, lap):
        """swe_height_correction (swe_time_step.cpp:558-604); returns the corrected lap."""
        lap = self._f(lap).copy()
        if lib().orc_swe_height_correction(self.h, _p(x), _p(y), _p(z), _pp([self._f(v) for v in vel]),
                                           _p(self._f(height)), _p(self._f(h_t)), dt, _p(lap)):
            raise RuntimeError("oracle swe_height_correction: geometry error")
        return lap

    def swe_back_rk4_step(self, x, y, z, area, vors, divs, height, dt, omega, eps, combine=0, topo=0, rule=0,
                          tracers=None):
        """swe_back_rk4_step (swe_time_step.cpp:606-811).  Returns (vors, divs, height,
        stage-1 velocity, stage-1 surface Laplacian, tracers)."""
        v, d, hh = [self._f(a).copy() for a in (vors, divs, height)]
        ntr = 0 if tracers is None else tracers.shape[-1]
        tr = np.zeros((self.Pl, self.m2, 1)) if tracers is None else np.ascontiguousarray(tracers, dtype=np.float64).copy()
        vel, lap = self._zp(3), self._zp()
        err = lib().orc_swe_back_rk4_step(self.h, _p(x), _p(y), _p(z), _p(area), _p(v), _p(d), _p(hh), _p(tr), ntr,
                                          dt, omega, eps, combine, topo, rule, _pp(vel), _p(lap))
        if err:
            raise RuntimeError(f"oracle swe_back_rk4_step: error {err}")
        return v, d, hh, vel, lap, (tr if ntr else None)

    def interp_points(self, x, y, z, fields):
        """latlon_interp_host at arbitrary points (1-D x, y, z); fields [Pl][m2] each."""
        x, y, z = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x, y, z)]
        fields = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
        out = [np.zeros(x.size) for _ in fields]
        err = lib().orc_interp_points(self.h, x.size, _p(x), _p(y), _p(z), len(fields), _pp(fields), _pp(out))
        if err:
            raise RuntimeError("oracle interp_points: geometry error")
        return out


def xyz_vel_to_uv_vel(x, y, z, vx, vy, vz):
    a = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x, y, z, vx, vy, vz)]
    u, v = np.zeros(a[0].size), np.zeros(a[0].size)
    lib().orc_xyz_vel_to_uv_vel(a[0].size, *[_p(t) for t in a], _p(u), _p(v))
    return u, v


def _grid_arrays(nlat, nlon, *arrs):
    out = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in arrs]
    for a in out:
        if a.size != nlat * nlon:
            raise ValueError(f"lat-lon field of {a.size} values, expected {nlat} x {nlon}")
    return out


def ll_tendencies(nlat, nlon, dx, omega, x, y, z, vor, div, h, topo, vel):
    """swe_tendency_computation_ll: (vor_tend, div_tend, height_tend), each [nlat*nlon]."""
    a = _grid_arrays(nlat, nlon, x, y, z, vor, div, h, topo, *vel)
    out = [np.zeros(nlat * nlon) for _ in range(3)]
    if lib().orc_ll_tendencies(nlat, nlon, float(dx), float(omega), *[_p(t) for t in a + out]):
        raise ValueError("ll_tendencies: grid under 3 x 3")
    return out


def ll_polar_filter(nlat, nlon, dx, x, y, z, f):
    """polar_filtering with zero-initialised coefficients; returns the filtered copy."""
    a = _grid_arrays(nlat, nlon, x, y, z)
    g = np.array(f, dtype=np.float64).ravel().copy()
    lib().orc_ll_polar_filter(nlat, nlon, float(dx), *[_p(t) for t in a], _p(g))
    return g


def ll_arrival_values(nlat, nlon, dx, dt, x, y, z, vor, div, h, vel, vor_t, div_t):
    """swe_compute_arrival_values_ll: (a_vor, a_div, a_height)."""
    a = _grid_arrays(nlat, nlon, x, y, z, vor, div, h, *vel, vor_t, div_t)
    out = [np.zeros(nlat * nlon) for _ in range(3)]
    if lib().orc_ll_arrival_values(nlat, nlon, float(dx), float(dt), *[_p(t) for t in a + out]):
        raise ValueError("ll_arrival_values: grid under 3 x 3")
    return out


def ll_field_averaging(nlat, nlon, x, y, z, f, epsilon):
    a = _grid_arrays(nlat, nlon, x, y, z, f)
    out = np.zeros(nlat * nlon)
    lib().orc_ll_field_averaging(nlat, nlon, *[_p(t) for t in a], _p(out), float(epsilon))
    return out


def ll_laplacian_diffusion(nlat, nlon, dx, dt, x, y, z, f):
    a = _grid_arrays(nlat, nlon, x, y, z, f)
    out = np.zeros(nlat * nlon)
    lib().orc_ll_laplacian_diffusion(nlat, nlon, float(dx), float(dt), *[_p(t) for t in a], _p(out))
    return out


def dedupe(x, y, z):
    """cubed_sphere_2d_to_1d, the reference's O(N^2) scan: (map, counts, x1d, y1d, z1d)."""
    x, y, z = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x, y, z)]
    n = x.size
    m = np.zeros(n, dtype=np.int32)
    c = np.zeros(n, dtype=np.int32)
    o = [np.zeros(n) for _ in range(3)]
    k = lib().orc_dedupe(n, _p(x), _p(y), _p(z), m.ctypes.data_as(_ip), c.ctypes.data_as(_ip), *[_p(t) for t in o])
    return m, c[:k].copy(), o[0][:k].copy(), o[1][:k].copy(), o[2][:k].copy()


def unify_boundary_vals(two_d_to_1d, counts, vals):
    """unify_boundary_vals (initialize_cubed_sphere.cpp:328-370): every copy of a unique point
    takes the mean of the copies.  Returns the new array."""
    m = np.ascontiguousarray(two_d_to_1d, dtype=np.int32).ravel()
    c = np.ascontiguousarray(counts, dtype=np.int32).ravel()
    v = np.array(vals, dtype=np.float64).ravel()
    rc = lib().orc_unify_boundary_vals(m.size, m.ctypes.data_as(_ip), c.ctypes.data_as(_ip), c.size, _p(v))
    if rc:
        raise ValueError("unify_boundary_vals: map index out of range")
    return v.reshape(np.shape(vals))


class OracleIrreg:
    """The irregular particle tree of the lat-lon path and its _ll passes
    (initialize_cubed_sphere.cpp:446-592, fmm_funcs.cpp:125-231, 394-419,
    swe_vel_impl.hpp:149-288), restated sequentially."""

    def __init__(self, x, y, z, leaf_size, degree, theta=0.7):
        self.x, self.y, self.z = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x, y, z)]
        self.N = self.x.size
        self.degree, self.n = degree, (degree + 1) ** 2
        self.h = lib().orc_irreg_create(self.N, _p(self.x), _p(self.y), _p(self.z), leaf_size, degree, theta)
        s = np.zeros(3, dtype=np.int64)
        lib().orc_irreg_sizes(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        self.P, self.levels, self.I = [int(v) for v in s]

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib().orc_irreg_destroy(self.h)
            self.h = None

    def panels(self):
        ints = np.zeros((self.P, 9), dtype=np.int32)
        dbls = np.zeros((self.P, 5))
        pts = np.zeros(self.N, dtype=np.int32)
        lib().orc_irreg_panels(self.h, ints.ctypes.data_as(_ip), _p(dbls), pts.ctypes.data_as(_ip))
        return ints, dbls, pts

    def interactions(self):
        out = np.zeros((self.I, 3), dtype=np.int32)
        lib().orc_irreg_interactions(self.h, out.ctypes.data_as(_ip))
        return out

    def upward_pass(self, area, pots):
        q = np.zeros((self.P, self.n))
        a, p = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (area, pots)]
        if lib().orc_irreg_upward_pass(self.h, _p(a), _p(p), _p(q)):
            raise RuntimeError("oracle irreg upward pass: geometry error")
        return q

    def downward_pass_3(self, proxy, sols=None):
        proxy = [np.ascontiguousarray(v, dtype=np.float64).copy() for v in proxy]
        sols = [np.zeros(self.N) for _ in range(3)] if sols is None else [np.array(v, dtype=np.float64) for v in sols]
        if lib().orc_irreg_downward_pass_3(self.h, _pp(proxy), _pp(sols)):
            raise RuntimeError("oracle irreg downward pass: geometry error")
        return proxy, sols

    def swe_vel_interactions(self, proxy_vors, proxy_divs, vors, divs, area, eps):
        tp = [np.zeros((self.P, self.n)) for _ in range(3)]
        vel = [np.zeros(self.N) for _ in range(3)]
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (proxy_vors, proxy_divs)]
        b = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (vors, divs, area)]
        lib().orc_irreg_swe_vel_interactions(self.h, _pp(tp), _pp(vel), _p(a[0]), _p(a[1]), _p(b[0]), _p(b[1]),
                                             _p(b[2]), eps)
        return tp, vel

    def swe_velocity(self, area, vors, divs, eps):
        """The lat-lon SWE velocity sequence: upward passes of both charges, interactions,
        downward pass (swe_time_step_ll.cpp's use of the _ll functors)."""
        qv = self.upward_pass(area, vors)
        qd = self.upward_pass(area, divs)
        tp, vel = self.swe_vel_interactions(qv, qd, vors, divs, area, eps)
        _, vel = self.downward_pass_3(tp, vel)
        return vel

    def swe_vel_tend_ll(self, nlat, nlon, dx, omega, area, vors, divs, height, topo, eps):
        """swe_vel_tend_computation (swe_time_step_ll.cpp:687-735): the FMM velocity, the
        tendencies and the polar filter of the height tendency.  Returns (vel, vor_t,
        div_t, h_t)."""
        vel = self.swe_velocity(area, vors, divs, eps)
        vt, dt_, ht = ll_tendencies(nlat, nlon, dx, omega, self.x, self.y, self.z, vors, divs, height, topo, vel)
        ht = ll_polar_filter(nlat, nlon, dx, self.x, self.y, self.z, ht)
        return vel, vt, dt_, ht

    def swe_back_rk4_step_ll(self, nlat, nlon, dx, dt, omega, area, vors, divs, height, topo, eps):
        """swe_back_rk4_step_ll (swe_time_step_ll.cpp:737-774): one semi-Lagrangian step
        on the lat-lon grid.  Returns (vors, divs, height, vel) -- vel of the start state."""
        vel, vt, dvt, _ = self.swe_vel_tend_ll(nlat, nlon, dx, omega, area, vors, divs, height, topo, eps)
        av, ad, ah = ll_arrival_values(nlat, nlon, dx, dt, self.x, self.y, self.z, vors, divs, height, vel, vt, dvt)
        d2 = ll_field_averaging(nlat, nlon, self.x, self.y, self.z, ad, 0.16)
        d3 = ll_laplacian_diffusion(nlat, nlon, dx, dt, self.x, self.y, self.z, d2)
        d3 = ll_polar_filter(nlat, nlon, dx, self.x, self.y, self.z, d3)
        return av, d3, ah, vel
