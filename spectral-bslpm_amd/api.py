"""Python host layer over libsbslpm.so, mirroring the reference's operator API.

Function names and argument meaning follow the reference's free functions
(fmm_funcs.hpp:16-54, bve_vel.hpp:10-14, inverse_poisson.hpp:10-12,
direct_sum.hpp:8-11, bve_time_step.hpp:7-13); arrays are torch tensors on the
GPU (float64, contiguous, [Pl][m2] for particles, [P][n] for proxies).  Errors
raise SbslpmError instead of aborting like Kokkos::abort / std::runtime_error.
All work runs on the context's HIP stream, which defaults to torch's current
stream so torch and library work are ordered.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import Params, SbslpmError, load

WEIGHTS = {"reference": 0, "clenshaw_curtis": 1, "equal_area": 2}
COMBINE = {"reference": 0, "textbook": 1}
TOPOGRAPHY = {"none": 0, "cone": 1, "gaussian": 2}
SWE_FORM = {"reg_laplacian": 1, "surface_correction": 2}
SWE_RULE = {"reference": 0, "leaf": 1}
PRECISION = {"fp64": 0, "mixed": 1, "mixed_valu": 2}


@dataclass
class RunConfig:
    """Hot-path subset of the reference RunConfig (run_config.hpp:6-44)."""
    levels: int
    interp_degree: int
    leaf_degree: int = None          # particle grid degree per leaf; None -> interp_degree (reference layout)
    fmm_theta: float = 0.7
    kernel_eps: float = 1e-4         # namelist.txt:16
    omega: float = 7.292e-5
    delta_t: float = 3600.0
    weights: str = "reference"       # reference | clenshaw_curtis | equal_area
    rk4_combine: str = "reference"   # reference | textbook
    precision: str = "fp64"          # fp64 | mixed (matrix cores) | mixed_valu
    tracer_count: int = 0
    use_graph: bool = False
    multi_rhs: bool = False          # ensembles: first stage as one multi-RHS product (mixed only)
    bf16_m2l: bool = False           # mixed: bf16 kernel-product operands for proxy-to-proxy sets
    topography: str = "none"         # RunConfig::topo_type: none | cone | gaussian
    swe_rule: str = "reference"      # reference | leaf (shallow-water step indexing, see sbslpm.h)
    gravity: float = 0.0             # free-surface Laplacian coefficient (g/a^2 in metres); 0: the reference's 1
    mountain: tuple = (0.0, 1.0, 0.0, 0.0)  # gaussian topography: height, width (rad), centre lat, lon
    swe_form: tuple = ()             # subset of SWE_FORM: departures from swe_back_rk4_step as written

    def to_params(self):
        p = Params()
        p.levels = self.levels
        p.interp_degree = self.interp_degree
        p.leaf_degree = self.interp_degree if self.leaf_degree is None else self.leaf_degree
        p.weight_rule = WEIGHTS[self.weights]
        p.rk4_combine = COMBINE[self.rk4_combine]
        p.precision = PRECISION[self.precision]
        p.tracer_count = self.tracer_count
        p.flags = (1 if self.use_graph else 0) | (2 if self.multi_rhs else 0) | (4 if self.bf16_m2l else 0)
        p.theta = self.fmm_theta
        p.kernel_eps = self.kernel_eps
        p.omega = self.omega
        p.delta_t = self.delta_t
        p.topography = TOPOGRAPHY[self.topography]
        p.swe_rule = SWE_RULE[self.swe_rule]
        p.gravity = self.gravity
        p.mountain_height, p.mountain_width, p.mountain_lat, p.mountain_lon = [float(v) for v in self.mountain]
        p.swe_form = sum(SWE_FORM[f] for f in self.swe_form)
        return p


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Sizes:
    def __init__(self, s):
        self.P, self.Pl, self.n, self.m2, self.I, self.N, self.offset, self.tiles = [int(v) for v in s]


class HostTree:
    """Host-side tree, particle grid and interaction list (no GPU needed)."""

    def __init__(self, cfg: RunConfig):
        self._lib = load()
        self.cfg = cfg
        h = ctypes.c_void_p()
        p = cfg.to_params()
        rc = self._lib.sbslpm_tree_build(ctypes.byref(p), ctypes.byref(h))
        if rc:
            raise SbslpmError(rc, "tree_build: invalid parameters")
        self.h = h
        s = np.zeros(8, dtype=np.int64)
        self._lib.sbslpm_tree_sizes(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        self.sizes = Sizes(s)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.sbslpm_tree_destroy(self.h)
            self.h = None

    def spec_points(self):
        s = self.sizes
        out = [np.zeros((s.Pl, s.m2)) for _ in range(4)]
        self._lib.sbslpm_tree_spec_points(self.h, *[_dp(a) for a in out])
        return out

    def interaction_list(self):
        out = np.zeros((self.sizes.I, 3), dtype=np.int32)
        self._lib.sbslpm_tree_interaction_list(self.h, _ip(out))
        return out

    def panels(self):
        ints = np.zeros((self.sizes.P, 8), dtype=np.int32)
        dbls = np.zeros((self.sizes.P, 6))
        self._lib.sbslpm_tree_panels(self.h, _ip(ints), _dp(dbls))
        return ints, dbls

    def locate(self, x, y, z):
        return self._lib.sbslpm_tree_locate(self.h, float(x), float(y), float(z))

    def transfer_matrices(self):
        d1 = self.cfg.interp_degree + 1
        m = (self.cfg.interp_degree if self.cfg.leaf_degree is None else self.cfg.leaf_degree) + 1
        hi, lo, p2m = np.zeros((d1, d1)), np.zeros((d1, d1)), np.zeros((d1, m))
        self._lib.sbslpm_tree_transfer_matrices(self.h, _dp(hi), _dp(lo), _dp(p2m))
        return hi, lo, p2m

    def dedupe(self):
        """cubed_sphere_2d_to_1d (initialize_cubed_sphere.cpp:290-324): (two_d_to_1d [Pl, m2],
        counts [n1d], x1d, y1d, z1d [n1d]) of the unique particle positions."""
        s = self.sizes
        m = np.zeros((s.Pl, s.m2), dtype=np.int32)
        cnt = np.zeros(s.N, dtype=np.int32)
        xyz = [np.zeros(s.N) for _ in range(3)]
        n1d = np.zeros(1, dtype=np.int64)
        rc = self._lib.sbslpm_tree_dedupe(self.h, _ip(m), _ip(cnt), *[_dp(a) for a in xyz],
                                          n1d.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        if rc:
            raise SbslpmError(rc, "tree_dedupe")
        k = int(n1d[0])
        return m, cnt[:k].copy(), xyz[0][:k].copy(), xyz[1][:k].copy(), xyz[2][:k].copy()


def cubed_sphere_midpoints(levels):
    """Cell centres and areas of the equiangular grid with 2^levels cells per face edge
    (initialize_cubed_sphere.cpp:10-62): x, y, z, area, each [6 * 4^levels]."""
    n = 6 * 4 ** levels
    out = [np.zeros(n) for _ in range(4)]
    rc = load().sbslpm_cubed_sphere_midpoints(int(levels), *[_dp(a) for a in out])
    if rc:
        raise SbslpmError(rc, "cubed_sphere_midpoints: levels must be in 0..14")
    return out


class HostIrregTree:
    """Host view of the irregular particle tree (initialize_cubed_sphere.cpp:446-592,
    fmm_funcs.cpp:125-231); points x, y, z are flat arrays in any order."""

    def __init__(self, x, y, z, leaf_size, interp_degree, theta=0.7):
        self._lib = load()
        a = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x, y, z)]
        h = ctypes.c_void_p()
        rc = self._lib.sbslpm_irreg_tree_build(a[0].size, *[_dp(v) for v in a], int(leaf_size), int(interp_degree),
                                               float(theta), ctypes.byref(h))
        if rc:
            raise SbslpmError(rc, "irreg_tree_build: invalid parameters")
        self.h = h
        s = np.zeros(6, dtype=np.int64)
        self._lib.sbslpm_irreg_tree_sizes(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        self.P, self.levels, self.n, self.N, self.I, self.tiles = [int(v) for v in s]

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.sbslpm_irreg_tree_destroy(self.h)
            self.h = None

    def panels(self):
        ints = np.zeros((self.P, 10), dtype=np.int32)
        dbls = np.zeros((self.P, 5))
        self._lib.sbslpm_irreg_tree_panels(self.h, _ip(ints), _dp(dbls))
        return ints, dbls

    def points(self):
        perm = np.zeros(self.N, dtype=np.int32)
        self._lib.sbslpm_irreg_tree_points(self.h, _ip(perm))
        return perm

    def interaction_list(self):
        out = np.zeros((self.I, 3), dtype=np.int32)
        self._lib.sbslpm_irreg_tree_interaction_list(self.h, _ip(out))
        return out


def _on_gpu(t):
    return t.is_cuda


def _ptr(t, count=None, what="array"):
    """Device pointer of a contiguous float64 GPU tensor (None -> NULL).

    With `count` the tensor must hold exactly that many elements.  Anything
    else (a host tensor, another dtype, a strided view, a short array) raises
    ValueError before the library sees the pointer, so no kernel reads or
    writes outside a caller's buffer.
    """
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: expected a torch tensor, got {type(t).__name__}")
    if count is not None and t.numel() != count:
        raise ValueError(f"{what}: expected {count} elements, got {t.numel()} (shape {tuple(t.shape)})")
    if t.dtype != torch.float64:
        raise ValueError(f"{what}: expected float64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous tensor")
    if not _on_gpu(t):
        raise ValueError(f"{what}: expected a tensor on the GPU, got {t.device}")
    return ctypes.c_void_p(t.data_ptr())


def _iptr(t, count=None, what="index array"):
    """Device pointer of a contiguous int32 GPU tensor, checked like _ptr."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: expected a torch tensor, got {type(t).__name__}")
    if count is not None and t.numel() != count:
        raise ValueError(f"{what}: expected {count} elements, got {t.numel()} (shape {tuple(t.shape)})")
    if not (t.dtype == torch.int32 and t.is_contiguous() and _on_gpu(t)):
        raise ValueError(f"{what}: expected a contiguous int32 tensor on the GPU")
    return ctypes.c_void_p(t.data_ptr())


def _seq(ts, k, what):
    """ts as a list of exactly k entries (tensors or None)."""
    ts = list(ts)
    if len(ts) != k:
        raise ValueError(f"{what}: expected {k} arrays, got {len(ts)}")
    return ts


def _ptr_array(ptrs):
    """A C array of device pointers (const double* const*) as one void* argument."""
    arr = (ctypes.c_void_p * len(ptrs))(*[p.value for p in ptrs])
    return arr, ctypes.cast(arr, ctypes.c_void_p)


class Context:
    """Device context: tree, tables, interaction lists and workspace on one GPU."""

    def __init__(self, cfg: RunConfig, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("sbslpm.Context needs a GPU (HIP device)")
        self._lib = load()
        self.cfg = cfg
        self.torch = torch
        h = ctypes.c_void_p()
        p = cfg.to_params()
        rc = self._lib.sbslpm_ctx_create(ctypes.byref(p), ctypes.byref(h))
        if rc:
            raise SbslpmError(rc, "ctx_create failed")
        self.h = h
        s = np.zeros(8, dtype=np.int64)
        self._lib.sbslpm_get_sizes(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        self.sizes = Sizes(s)
        self.set_stream(stream if stream is not None else torch.cuda.current_stream())

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.sbslpm_ctx_destroy(self.h)
            self.h = None

    # ---- plumbing
    def _check(self, rc, what):
        if rc:
            msg = self._lib.sbslpm_last_error(self.h)
            raise SbslpmError(rc, f"{what}: {msg.decode() if msg else ''}")

    # Argument checks.  Particle arrays hold Pl * m2 values per member, proxy
    # arrays P * n; every pointer handed to the library goes through one of these.
    def _p(self, t, k=1, what="particle array"):
        return _ptr(t, k * self.sizes.Pl * self.sizes.m2, what)

    def _x(self, t, what="proxy array"):
        return _ptr(t, self.sizes.P * self.sizes.n, what)

    def _p3(self, ts, what, k=1):
        return [self._p(t, k, what) for t in _seq(ts, 3, what)]

    def _x3(self, ts, what):
        return [self._x(t, what) for t in _seq(ts, 3, what)]

    def _opt3(self, ts, what, proxy=False):
        """Optional three-component array (displacements): None or three tensors."""
        if ts is None:
            return [None] * 3
        return self._x3(ts, what) if proxy else self._p3(ts, what)

    def _stages(self, vels, what):
        """The 12 stage velocities [v0, v1, v2, v3] (three tensors each) as a pointer array."""
        flat = [t for stage in _seq(vels, 4, what) for t in _seq(stage, 3, what)]
        return _ptr_array([self._p(t, what=what) for t in flat])

    def set_stream(self, stream):
        self.stream = stream
        self._check(self._lib.sbslpm_ctx_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)), "set_stream")

    def set_precision(self, precision):
        self._check(self._lib.sbslpm_ctx_set_precision(self.h, PRECISION[precision]), "set_precision")
        self.cfg.precision = precision

    KERNELS = {"bve_fp64": 0, "bve_mixed": 1, "log_fp64": 2, "log_mixed": 3, "swe_fp64": 4, "lap_fp64": 5,
               "swe_mixed": 6, "lap_mixed": 7, "bve_mrhs": 8}

    def enable_timing(self, on=True):
        self._check(self._lib.sbslpm_ctx_enable_timing(self.h, 1 if on else 0), "enable_timing")

    def kernel_time(self, kernel, reset=True):
        """(summed milliseconds, launch count) of one pair-kernel class since the last reset."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._lib.sbslpm_kernel_time(self.h, self.KERNELS[kernel], ctypes.byref(ms), ctypes.byref(n),
                                                 1 if reset else 0), "kernel_time")
        return ms.value, n.value

    def synchronize(self):
        self._check(self._lib.sbslpm_synchronize(self.h), "synchronize")

    def zeros_part(self, k=None):
        shape = (self.sizes.Pl, self.sizes.m2)
        if k is None:
            return self.torch.zeros(shape, dtype=self.torch.float64, device="cuda")
        return [self.torch.zeros(shape, dtype=self.torch.float64, device="cuda") for _ in range(k)]

    def zeros_proxy(self, k=None):
        shape = (self.sizes.P, self.sizes.n)
        if k is None:
            return self.torch.zeros(shape, dtype=self.torch.float64, device="cuda")
        return [self.torch.zeros(shape, dtype=self.torch.float64, device="cuda") for _ in range(k)]

    def spec_points(self):
        s = self.sizes
        out = [np.zeros((s.Pl, s.m2)) for _ in range(4)]
        self._check(self._lib.sbslpm_spec_points(self.h, *[_dp(a) for a in out]), "spec_points")
        return out

    def interaction_list(self):
        out = np.zeros((self.sizes.I, 3), dtype=np.int32)
        self._check(self._lib.sbslpm_interaction_list(self.h, _ip(out)), "interaction_list")
        return out

    # ---- reference operator API
    def upward_pass(self, area, pots, q_part=None, q_proxy=None):
        """upward_pass, fmm_funcs.cpp:341-355."""
        q_part = self.zeros_part() if q_part is None else q_part
        q_proxy = self.zeros_proxy() if q_proxy is None else q_proxy
        self._check(self._lib.sbslpm_upward_pass(self.h, self._p(area, what="area"), self._p(pots, what="pots"),
                                                 self._p(q_part, what="q_part"), self._x(q_proxy, what="q_proxy")),
                    "upward_pass")
        return q_part, q_proxy

    def bve_vel_interactions(self, q_part, q_proxy, disp_part=None, disp_proxy=None, phi_part=None,
                             phi_proxy=None):
        """bve_vel_interactions, bve_vel.cpp:9-37 (accumulates)."""
        phi_part = self.zeros_part(3) if phi_part is None else phi_part
        phi_proxy = self.zeros_proxy(3) if phi_proxy is None else phi_proxy
        args = (self._opt3(disp_part, "disp_part") + self._opt3(disp_proxy, "disp_proxy", proxy=True)
                + [self._p(q_part, what="q_part"), self._x(q_proxy, what="q_proxy")]
                + self._p3(phi_part, "phi_part") + self._x3(phi_proxy, "phi_proxy"))
        self._check(self._lib.sbslpm_bve_vel_interactions(self.h, *args), "bve_vel_interactions")
        return phi_part, phi_proxy

    def swe_vel_interactions(self, qz_part, qz_proxy, qd_part, qd_proxy, disp_part=None, disp_proxy=None,
                             phi_part=None, phi_proxy=None, variant=1):
        """swe_vel_interactions (variant 1, swe_vel.cpp:9-35) or swe_vel_interactions_2 (variant 2,
        swe_vel.cpp:37-67); accumulates."""
        phi_part = self.zeros_part(3) if phi_part is None else phi_part
        phi_proxy = self.zeros_proxy(3) if phi_proxy is None else phi_proxy
        args = (self._opt3(disp_part, "disp_part") + self._opt3(disp_proxy, "disp_proxy", proxy=True)
                + [self._p(qz_part, what="qz_part"), self._x(qz_proxy, what="qz_proxy"),
                   self._p(qd_part, what="qd_part"), self._x(qd_proxy, what="qd_proxy")]
                + self._p3(phi_part, "phi_part") + self._x3(phi_proxy, "phi_proxy"))
        fn = self._lib.sbslpm_swe_vel_interactions if variant == 1 else self._lib.sbslpm_swe_vel_interactions_2
        self._check(fn(self.h, *args), "swe_vel_interactions")
        return phi_part, phi_proxy

    def swe_velocity(self, area, vor, div, disp_part=None, disp_proxy=None, vel=None, variant=1):
        """Upward passes of vor and div + SWE interactions + downward pass (swe_time_step.cpp:640-653)."""
        vel = self.zeros_part(3) if vel is None else vel
        args = ([self._p(area, what="area"), self._p(vor, what="vor"), self._p(div, what="div")]
                + self._opt3(disp_part, "disp_part") + self._opt3(disp_proxy, "disp_proxy", proxy=True)
                + self._p3(vel, "vel"))
        self._check(self._lib.sbslpm_swe_velocity(self.h, variant, *args), "swe_velocity")
        return vel

    def poisson_fmm_interactions(self, q_part, q_proxy, psi_part=None, psi_proxy=None):
        """poisson_fmm_interactions, inverse_poisson.cpp:9-33 (accumulates)."""
        psi_part = self.zeros_part() if psi_part is None else psi_part
        psi_proxy = self.zeros_proxy() if psi_proxy is None else psi_proxy
        self._check(self._lib.sbslpm_poisson_fmm_interactions(
            self.h, self._p(q_part, what="q_part"), self._x(q_proxy, what="q_proxy"),
            self._p(psi_part, what="psi_part"), self._x(psi_proxy, what="psi_proxy")), "poisson_fmm_interactions")
        return psi_part, psi_proxy

    def downward_pass(self, phi_proxy, phi_part=None, sols=None):
        """downward_pass, fmm_funcs.cpp:357-372."""
        sols = self.zeros_part() if sols is None else sols
        self._check(self._lib.sbslpm_downward_pass(self.h, self._x(phi_proxy, what="phi_proxy"),
                                                   self._p(phi_part, what="phi_part"), self._p(sols, what="sols")),
                    "downward_pass")
        return sols

    def downward_pass_3(self, phi_proxy, phi_part=None, sols=None):
        """downward_pass_3, fmm_funcs.cpp:374-392."""
        sols = self.zeros_part(3) if sols is None else sols
        args = self._x3(phi_proxy, "phi_proxy") + self._opt3(phi_part, "phi_part") + self._p3(sols, "sols")
        self._check(self._lib.sbslpm_downward_pass_3(self.h, *args), "downward_pass_3")
        return sols

    def bve_velocity(self, area, vor, disp_part=None, disp_proxy=None, vel=None):
        """Upward pass + interactions + downward pass (bve_time_step.cpp:321-344)."""
        vel = self.zeros_part(3) if vel is None else vel
        args = ([self._p(area, what="area"), self._p(vor, what="vor")] + self._opt3(disp_part, "disp_part")
                + self._opt3(disp_proxy, "disp_proxy", proxy=True) + self._p3(vel, "vel"))
        self._check(self._lib.sbslpm_bve_velocity(self.h, *args), "bve_velocity")
        return vel

    def poisson_solve(self, area, f, psi=None):
        """poisson_eq_fmm.cpp:129-147."""
        psi = self.zeros_part() if psi is None else psi
        self._check(self._lib.sbslpm_poisson_solve(self.h, self._p(area, what="area"), self._p(f, what="f"),
                                                   self._p(psi, what="psi")), "poisson_solve")
        return psi

    def poisson_solve_targets(self, area, f, tx, ty, tz, psi=None):
        """Poisson solve evaluated at arbitrary unit-sphere targets (1-D arrays):
        leaf far field by BLI (child_panel_interp_3_ll, fmm_funcs_impl.hpp:318-364)
        plus the leaf's near field."""
        nt = tx.numel()
        psi = self.torch.zeros(nt, dtype=self.torch.float64, device="cuda") if psi is None else psi
        t3 = [_ptr(t, nt, "targets") for t in (tx, ty, tz)]
        self._check(self._lib.sbslpm_poisson_solve_targets(self.h, self._p(area, what="area"), self._p(f, what="f"),
                                                           nt, *t3, _ptr(psi, nt, "psi")), "poisson_solve_targets")
        return psi

    def bve_velocity_targets(self, area, vor, tx, ty, tz, vel=None):
        """Biot-Savart velocity at arbitrary unit-sphere targets (1-D arrays)."""
        nt = tx.numel()
        vel = [self.torch.zeros(nt, dtype=self.torch.float64, device="cuda") for _ in range(3)] if vel is None else vel
        t3 = [_ptr(t, nt, "targets") for t in (tx, ty, tz)]
        v3 = [_ptr(v, nt, "vel") for v in _seq(vel, 3, "vel")]
        self._check(self._lib.sbslpm_bve_velocity_targets(self.h, self._p(area, what="area"), self._p(vor, what="vor"),
                                                          nt, *t3, *v3), "bve_velocity_targets")
        return vel

    def direct_sum_inv_lap(self, x, y, z, area, pots, soln=None, lb=0, ub=None):
        """direct_sum_inv_lap, direct_sum.cpp:8-15 (1-D arrays, accumulates)."""
        n = x.numel()
        ub = n if ub is None else ub
        soln = self.torch.zeros(n, dtype=self.torch.float64, device="cuda") if soln is None else soln
        a = [_ptr(t, n, "direct_sum_inv_lap input") for t in (x, y, z, area, pots, soln)]
        self._check(self._lib.sbslpm_direct_sum_inv_lap(self.h, n, *a, lb, ub), "direct_sum_inv_lap")
        return soln

    def direct_sum_inv_lap_targets(self, tx, ty, tz, sx, sy, sz, area, pots, soln=None):
        nt, ns = tx.numel(), sx.numel()
        soln = self.torch.zeros(nt, dtype=self.torch.float64, device="cuda") if soln is None else soln
        t3 = [_ptr(t, nt, "targets") for t in (tx, ty, tz)]
        s5 = [_ptr(t, ns, "sources") for t in (sx, sy, sz, area, pots)]
        self._check(self._lib.sbslpm_direct_sum_inv_lap_targets(self.h, nt, *t3, ns, *s5, _ptr(soln, nt, "soln")),
                    "direct_sum_inv_lap_targets")
        return soln

    def direct_sum_bve_vel_targets(self, tx, ty, tz, sx, sy, sz, area, vor, vel=None):
        """Biot-Savart direct sum at separate targets over undisplaced sources (the
        context's eps); vel (three 1-D tensors) accumulated."""
        nt, ns = tx.numel(), sx.numel()
        vel = [self.torch.zeros(nt, dtype=self.torch.float64, device="cuda") for _ in range(3)] if vel is None else vel
        t3 = [_ptr(t, nt, "targets") for t in (tx, ty, tz)]
        s5 = [_ptr(t, ns, "sources") for t in (sx, sy, sz, area, vor)]
        v3 = [_ptr(v, nt, "vel") for v in _seq(vel, 3, "vel")]
        self._check(self._lib.sbslpm_direct_sum_bve_vel_targets(self.h, nt, *t3, ns, *s5, *v3),
                    "direct_sum_bve_vel_targets")
        return vel

    def direct_sum_bve_vel(self, x, y, z, area, vor, disp=None, vel=None, lb=0, ub=None):
        n = x.numel()
        ub = n if ub is None else ub
        vel = [self.torch.zeros(n, dtype=self.torch.float64, device="cuda") for _ in range(3)] if vel is None else vel
        d = [None] * 3 if disp is None else [_ptr(t, n, "disp") for t in _seq(disp, 3, "disp")]
        a = [_ptr(t, n, "direct_sum_bve_vel input") for t in (x, y, z, area, vor)]
        v3 = [_ptr(v, n, "vel") for v in _seq(vel, 3, "vel")]
        self._check(self._lib.sbslpm_direct_sum_bve_vel(self.h, n, *a, *d, *v3, lb, ub), "direct_sum_bve_vel")
        return vel

    def bve_displacement(self, x, y, z, vel, vor, tau, new_vor=None, disp_part=None, disp_proxy=None):
        """bve_displacement_upward_pass, bve_time_step.cpp:162-170."""
        new_vor = self.zeros_part() if new_vor is None else new_vor
        disp_part = self.zeros_part(3) if disp_part is None else disp_part
        disp_proxy = self.zeros_proxy(3) if disp_proxy is None else disp_proxy
        a = [self._p(t, what="x, y, z") for t in (x, y, z)] + self._p3(vel, "vel") + [self._p(vor, what="vor")]
        b = [self._p(new_vor, what="new_vor")] + self._p3(disp_part, "disp_part") + self._x3(disp_proxy, "disp_proxy")
        self._check(self._lib.sbslpm_bve_displacement(self.h, *a, float(tau), *b), "bve_displacement")
        return new_vor, disp_part, disp_proxy

    def bve_find_departure_points(self, x, y, z, vels, dt, dep=None):
        """bve_find_departure_points, bve_time_step.cpp:172-309; vels = [v0, v1, v2, v3], each 3 tensors."""
        dep = self.zeros_part(3) if dep is None else dep
        _keep, arr = self._stages(vels, "vels")
        self._check(self._lib.sbslpm_bve_find_departure_points(
            self.h, *[self._p(t, what="x, y, z") for t in (x, y, z)], arr, float(dt), *self._p3(dep, "dep")),
            "departure_points")
        return dep

    def bve_departure_to_target(self, z, dep, vor, tracers=None, new_vor=None, new_tracers=None):
        """bve_departure_to_target, bve_time_step.cpp:15-67."""
        new_vor = self.zeros_part() if new_vor is None else new_vor
        ntr = 0 if tracers is None else tracers.shape[-1]
        if ntr and new_tracers is None:
            new_tracers = self.torch.zeros_like(tracers)
        tr = [self._p(tracers, ntr, "tracers"), self._p(new_tracers, ntr, "new_tracers")] if ntr else [None, None]
        self._check(self._lib.sbslpm_bve_departure_to_target(
            self.h, self._p(z, what="z"), *self._p3(dep, "dep"), self._p(vor, what="vor"), tr[0], ntr,
            self._p(new_vor, what="new_vor"), tr[1]), "departure_to_target")
        return new_vor, new_tracers

    def bve_back_rk4_step_ensemble(self, x, y, z, area, vors, vel, time=0.0):
        """One backward RK4 step of every member of an ensemble (vors [members, Pl, m2] updated in
        place, vel = 3 tensors [members, Pl, m2] receive the stage-0 velocities); all members share
        each kernel launch."""
        nmem = int(vors.shape[0]) if vors.dim() >= 2 else 0
        if nmem < 1:
            raise ValueError("vors: expected shape [members, Pl, m2]")
        args = ([self._p(t, what="x, y, z, area") for t in (x, y, z, area)] + [self._p(vors, nmem, "vors")]
                + self._p3(vel, "vel", k=nmem))
        self._check(self._lib.sbslpm_bve_back_rk4_step_ensemble(self.h, nmem, *args, float(time)),
                    "bve_back_rk4_step_ensemble")
        return vors

    def bve_velocity_adjoint(self, area, ubar, vor_bar=None):
        """Transpose of the undisplaced velocity evaluation (exact transpose of the FP64 FMM operator):
        vor_bar = K^T ubar (ubar = 3 tensors)."""
        vor_bar = self.zeros_part() if vor_bar is None else vor_bar
        self._check(self._lib.sbslpm_bve_velocity_adjoint(self.h, self._p(area, what="area"), *self._p3(ubar, "ubar"),
                                                          self._p(vor_bar, what="vor_bar")), "bve_velocity_adjoint")
        return vor_bar

    def autotune(self, area, vor):
        """Times the pair kernel's launch shapes on this context's interaction list and keeps the
        fastest; returns [(waves, xcd, ms per launch), ...] and the chosen (waves, xcd)."""
        rep = np.zeros((9, 3))
        n = ctypes.c_int(0)
        self._check(self._lib.sbslpm_autotune(self.h, self._p(area, what="area"), self._p(vor, what="vor"), _dp(rep),
                                              ctypes.byref(n)), "autotune")
        return [tuple(r) for r in rep[:n.value]], self.pair_shape()

    def autotune_ensemble(self, area, vors):
        """Times the multi-right-hand-side pair kernel of ensemble steps (mixed context with
        multi_rhs=True) over waves per workgroup x XCD mapping x chunks of four members per
        wave, with the charges of vors [members][Pl][m2], and keeps the fastest; returns
        [(waves, xcd, chunks, ms per launch), ...] and the chosen (waves, xcd, chunks)."""
        rep = np.zeros((18, 4))
        n = ctypes.c_int(0)
        nmem = int(vors.shape[0]) if vors.dim() >= 2 else 0
        if nmem < 1:
            raise ValueError("vors: expected shape [members, Pl, m2]")
        self._check(self._lib.sbslpm_autotune_ensemble(self.h, nmem, self._p(area, what="area"),
                                                       self._p(vors, nmem, "vors"), _dp(rep), ctypes.byref(n)),
                    "autotune_ensemble")
        rows = [tuple(r) for r in rep[:n.value]]
        best = min(rows, key=lambda r: r[3])
        return rows, (int(best[0]), int(best[1]), int(best[2]))

    def unify_boundary_vals(self, two_d_to_1d, counts, vals):
        """unify_boundary_vals (initialize_cubed_sphere.cpp:364-370): every copy of a unique
        node takes the mean of its copies, in place.  two_d_to_1d [Pl, m2] and counts [n1d]
        are int32 device tensors (from HostTree.dedupe); vals [Pl, m2] float64."""
        npart = self.sizes.Pl * self.sizes.m2
        self._check(self._lib.sbslpm_unify_boundary_vals(self.h, _iptr(two_d_to_1d, npart, "two_d_to_1d"),
                                                         _iptr(counts, what="counts"), int(counts.numel()),
                                                         self._p(vals, what="vals")), "unify_boundary_vals")
        return vals

    def set_pair_shape(self, waves, xcd):
        self._check(self._lib.sbslpm_set_pair_shape(self.h, int(waves), int(xcd)), "set_pair_shape")

    def pair_shape(self):
        w, x = ctypes.c_int(0), ctypes.c_int(0)
        self._lib.sbslpm_get_pair_shape(self.h, ctypes.byref(w), ctypes.byref(x))
        return w.value, x.value

    def irregular(self, x, y, z, leaf_size):
        """Irregular particle tree over host points x, y, z (flat, any order) using this
        context's stream, interp_degree, theta and kernel_eps."""
        return IrregContext(self, x, y, z, leaf_size)

    def interp_points(self, x, y, z, fields, out=None):
        """BLI of leaf fields [Pl, m2] (a list of up to 8 tensors) at the points (x, y, z) (1-D
        tensors of one length): latlon_interp_host, interp_funcs.hpp:22-60."""
        n = x.numel()
        fields = list(fields)
        out = [self.torch.zeros(n, dtype=self.torch.float64, device="cuda") for _ in fields] if out is None else out
        _kf, fa = _ptr_array([self._p(f, what="fields") for f in fields])
        _ko, oa = _ptr_array([_ptr(o, n, "out") for o in _seq(out, len(fields), "out")])
        self._check(self._lib.sbslpm_interp_points(self.h, n, *[_ptr(t, n, "points") for t in (x, y, z)],
                                                   len(fields), fa, oa), "interp_points")
        return out

    def interp_points_adjoint(self, x, y, z, out_bar, field_bar=None):
        """Transpose of interp_points: field_bar[q] [Pl, m2] += B^T out_bar[q]."""
        n = x.numel()
        out_bar = list(out_bar)
        field_bar = [self.zeros_part() for _ in out_bar] if field_bar is None else field_bar
        _ko, ob = _ptr_array([_ptr(o, n, "out_bar") for o in out_bar])
        _kf, fb = _ptr_array([self._p(f, what="field_bar") for f in _seq(field_bar, len(out_bar), "field_bar")])
        self._check(self._lib.sbslpm_interp_points_adjoint(self.h, n, *[_ptr(t, n, "points") for t in (x, y, z)],
                                                           len(out_bar), ob, fb), "interp_points_adjoint")
        return field_bar

    def xyz_vel_to_uv_vel(self, x, y, z, vel, u=None, v=None):
        """Zonal and meridional velocity, xyz_vel_to_uv_vel (general_utils_impl.hpp:195-229)."""
        vel = _seq(vel, 3, "vel")
        n = vel[0].numel()
        u = self.torch.zeros_like(vel[0]) if u is None else u
        v = self.torch.zeros_like(vel[0]) if v is None else v
        a = [_ptr(t, n, "xyz_vel_to_uv_vel input") for t in (x, y, z, *vel)]
        self._check(self._lib.sbslpm_xyz_vel_to_uv_vel(self.h, n, *a, _ptr(u, n, "u"), _ptr(v, n, "v")),
                    "xyz_vel_to_uv_vel")
        return u, v

    # ---- lat-lon shallow-water grid operators (swe_time_step_ll.cpp); grid = (lat_count,
    # lon_count, grid_spacing in degrees), fields [lat_count * lon_count] float64 on the GPU
    def _ll(self, grid, *ts, what):
        nlat, nlon, dx = int(grid[0]), int(grid[1]), float(grid[2])
        return (nlat, nlon, dx), [_ptr(t, nlat * nlon, what) for t in ts]

    def _ll_new(self, like, k):
        return [self.torch.empty_like(like) for _ in range(k)]

    def ll_tendencies(self, grid, x, y, z, vors, divs, height, topo, vel, omega=None, out=None):
        """swe_tendency_computation_ll: (vor_tend, div_tend, height_tend)."""
        from .latlon import OMEGA
        out = self._ll_new(vors, 3) if out is None else _seq(out, 3, "out")
        (nlat, nlon, dx), a = self._ll(grid, x, y, z, vors, divs, height, topo, *_seq(vel, 3, "vel"), *out,
                                       what="ll_tendencies array")
        self._check(self._lib.sbslpm_ll_tendencies(self.h, nlat, nlon, dx, float(OMEGA if omega is None else omega),
                                                   *a), "ll_tendencies")
        return out

    def ll_polar_filter(self, grid, x, y, z, field):
        """polar_filtering in place (zero-initialised coefficients)."""
        (nlat, nlon, dx), a = self._ll(grid, x, y, z, field, what="ll_polar_filter array")
        self._check(self._lib.sbslpm_ll_polar_filter(self.h, nlat, nlon, dx, *a), "ll_polar_filter")
        return field

    def ll_arrival_values(self, grid, delta_t, x, y, z, vors, divs, height, vel, vor_tend, div_tend, out=None):
        """swe_compute_arrival_values_ll: (a_vor, a_div, a_height)."""
        out = self._ll_new(vors, 3) if out is None else _seq(out, 3, "out")
        (nlat, nlon, dx), a = self._ll(grid, x, y, z, vors, divs, height, *_seq(vel, 3, "vel"), vor_tend, div_tend,
                                       *out, what="ll_arrival_values array")
        self._check(self._lib.sbslpm_ll_arrival_values(self.h, nlat, nlon, dx, float(delta_t), *a),
                    "ll_arrival_values")
        return out

    def ll_field_averaging(self, grid, x, y, z, field, epsilon, out=None):
        out = self.torch.empty_like(field) if out is None else out
        (nlat, nlon, _), a = self._ll(grid, x, y, z, field, out, what="ll_field_averaging array")
        self._check(self._lib.sbslpm_ll_field_averaging(self.h, nlat, nlon, *a, float(epsilon)), "ll_field_averaging")
        return out

    def ll_laplacian_diffusion(self, grid, dt, x, y, z, field, out=None):
        out = self.torch.empty_like(field) if out is None else out
        (nlat, nlon, dx), a = self._ll(grid, x, y, z, field, out, what="ll_laplacian_diffusion array")
        self._check(self._lib.sbslpm_ll_laplacian_diffusion(self.h, nlat, nlon, dx, float(dt), *a),
                    "ll_laplacian_diffusion")
        return out

    def bve_back_rk4_step_adjoint(self, x, y, z, area, vor, new_vor_bar, vor_bar=None):
        """Gradient of one bve_back_rk4_step with respect to the vorticity before it: vor_bar =
        (d new_vor / d vor)^T new_vor_bar at the state vor (not modified; the step's forward
        pass is recomputed inside)."""
        vor_bar = self.zeros_part() if vor_bar is None else vor_bar
        a = [self._p(t, what="bve_back_rk4_step_adjoint input") for t in (x, y, z, area, vor, new_vor_bar)]
        self._check(self._lib.sbslpm_bve_back_rk4_step_adjoint(self.h, *a, self._p(vor_bar, what="vor_bar")),
                    "bve_back_rk4_step_adjoint")
        return vor_bar

    def bve_back_rk4_step_record(self, x, y, z, area, vor, vor_out=None, stages=None):
        """One step keeping the stage velocities: returns (vor_out, stages [12, Pl, m2])."""
        vor_out = self.zeros_part() if vor_out is None else vor_out
        if stages is None:
            stages = self.torch.zeros((12,) + tuple(vor.shape), dtype=self.torch.float64, device="cuda")
        a = [self._p(t, what="bve_back_rk4_step_record input") for t in (x, y, z, area, vor)]
        self._check(self._lib.sbslpm_bve_back_rk4_step_record(self.h, *a, self._p(vor_out, what="vor_out"),
                                                              self._p(stages, 12, "stages")),
                    "bve_back_rk4_step_record")
        return vor_out, stages

    def bve_back_rk4_step_adjoint_stages(self, x, y, z, area, vor, stages, new_vor_bar, vor_bar=None):
        """bve_back_rk4_step_adjoint with the step's recorded stage velocities."""
        vor_bar = self.zeros_part() if vor_bar is None else vor_bar
        a = [self._p(t, what="bve_back_rk4_step_adjoint_stages input") for t in (x, y, z, area, vor)]
        self._check(self._lib.sbslpm_bve_back_rk4_step_adjoint_stages(
            self.h, *a, self._p(stages, 12, "stages"), self._p(new_vor_bar, what="new_vor_bar"),
            self._p(vor_bar, what="vor_bar")), "bve_back_rk4_step_adjoint_stages")
        return vor_bar

    def bve_departure_to_target_adjoint(self, dep, new_vor_bar, vor_bar=None):
        """Transpose of the remap interpolation: vor_bar += B^T new_vor_bar."""
        vor_bar = self.zeros_part() if vor_bar is None else vor_bar
        self._check(self._lib.sbslpm_bve_departure_to_target_adjoint(
            self.h, *self._p3(dep, "dep"), self._p(new_vor_bar, what="new_vor_bar"),
            self._p(vor_bar, what="vor_bar")), "bve_departure_to_target_adjoint")
        return vor_bar

    def swe_displacement(self, vel, tau, disp_part=None, disp_proxy=None):
        """swe_displacement_upward_pass, swe_time_step.cpp:265-276."""
        disp_part = self.zeros_part(3) if disp_part is None else disp_part
        disp_proxy = self.zeros_proxy(3) if disp_proxy is None else disp_proxy
        self._check(self._lib.sbslpm_swe_displacement(self.h, *self._p3(vel, "vel"), float(tau),
                                                      *self._p3(disp_part, "disp_part"),
                                                      *self._x3(disp_proxy, "disp_proxy")), "swe_displacement")
        return disp_part, disp_proxy

    # ---- shallow-water step (swe_time_step.cpp) ----
    def laplacian(self, f, lap=None):
        """laplacian (deriv_funcs.cpp:27-37): spectral surface Laplacian on each leaf."""
        lap = self.zeros_part() if lap is None else lap
        self._check(self._lib.sbslpm_laplacian(self.h, self._p(f, what="f"), self._p(lap, what="lap")), "laplacian")
        return lap

    def laplacian_interactions(self, q_part, q_proxy, disp_part=None, disp_proxy=None, phi_part=None,
                               phi_proxy=None):
        """laplacian_interactions (laplacian.cpp:6-35); accumulates."""
        phi_part = self.zeros_part() if phi_part is None else phi_part
        phi_proxy = self.zeros_proxy() if phi_proxy is None else phi_proxy
        self._check(self._lib.sbslpm_laplacian_interactions(
            self.h, *self._opt3(disp_part, "disp_part"), *self._opt3(disp_proxy, "disp_proxy", proxy=True),
            self._p(q_part, what="q_part"), self._x(q_proxy, what="q_proxy"), self._p(phi_part, what="phi_part"),
            self._x(phi_proxy, what="phi_proxy")), "laplacian_interactions")
        return phi_part, phi_proxy

    def apply_topography(self, x, y, z, height, disp_part=None, disp_proxy=None, eff=None):
        """apply_topography (topography.cpp:47-56) with the context's topography and swe_rule."""
        eff = self.zeros_part() if eff is None else eff
        self._check(self._lib.sbslpm_apply_topography(
            self.h, *[self._p(t, what="x, y, z") for t in (x, y, z)], *self._opt3(disp_part, "disp_part"),
            *self._opt3(disp_proxy, "disp_proxy", proxy=True), self._p(height, what="height"),
            self._p(eff, what="eff")), "apply_topography")
        return eff

    def swe_tendency(self, x, y, z, vel, vor, div, height, height_lap, out=None):
        """swe_tendency_computation (swe_time_step.cpp:76-161): (vor_t, div_t, h_t)."""
        out = self.zeros_part(3) if out is None else out
        a = ([self._p(t, what="x, y, z") for t in (x, y, z)] + self._p3(vel, "vel")
             + [self._p(t, what="vor, div, height, height_lap") for t in (vor, div, height, height_lap)])
        self._check(self._lib.sbslpm_swe_tendency(self.h, *a, *self._p3(out, "out")), "swe_tendency")
        return out

    def swe_height_correction(self, x, y, z, vel, height, h_t, dt, height_lap):
        """swe_height_correction (swe_time_step.cpp:558-604); height_lap updated in place."""
        a = ([self._p(t, what="x, y, z") for t in (x, y, z)] + self._p3(vel, "vel")
             + [self._p(height, what="height"), self._p(h_t, what="h_t")])
        self._check(self._lib.sbslpm_swe_height_correction(self.h, *a, float(dt),
                                                           self._p(height_lap, what="height_lap")),
                    "swe_height_correction")
        return height_lap

    def swe_back_rk4_step(self, x, y, z, area, vors, divs, height, vel=None, tracers=None, height_lap=None,
                          time=0.0):
        """swe_back_rk4_step (swe_time_step.cpp:606-811): vors, divs, height (and tracers) updated
        in place; returns (stage-1 velocity, stage-1 surface Laplacian)."""
        vel = self.zeros_part(3) if vel is None else vel
        height_lap = self.zeros_part() if height_lap is None else height_lap
        ntr = self.cfg.tracer_count
        tr = self._p(tracers, ntr, "tracers") if ntr > 0 else None
        a = [self._p(t, what="swe_back_rk4_step input") for t in (x, y, z, area, vors, divs, height)]
        self._check(self._lib.sbslpm_swe_back_rk4_step(self.h, *a, *self._p3(vel, "vel"), tr,
                                                       self._p(height_lap, what="height_lap"), float(time)),
                    "swe_back_rk4_step")
        return vel, height_lap

    def swe_back_rk4_step_record(self, x, y, z, area, vors, divs, height, out=None, stages=None):
        """One shallow-water step out of place, keeping what its reverse pass reads: returns
        (new_vors, new_divs, new_height, stages), stages [21][Pl][m2] = stage velocities V0..V3
        and tendencies T0..T2.  Needs tracer_count == 0."""
        out = [self.zeros_part() for _ in range(3)] if out is None else list(_seq(out, 3, "out"))
        if stages is None:
            stages = self.torch.zeros((21, self.sizes.Pl, self.sizes.m2), dtype=self.torch.float64, device="cuda")
        st = self._p(stages, 21, "stages")
        a = [self._p(t, what="swe_back_rk4_step_record input") for t in (x, y, z, area, vors, divs, height)]
        self._check(self._lib.sbslpm_swe_back_rk4_step_record(self.h, *a, *[self._p(t, what="out") for t in out], st),
                    "swe_back_rk4_step_record")
        return out[0], out[1], out[2], stages

    def swe_back_rk4_step_adjoint(self, x, y, z, area, vors, divs, height, bars, stages=None, out=None):
        """Gradient of one shallow-water step (swe_back_rk4_step, swe_time_step.cpp:606-811) with
        respect to (vors, divs, height) before it, for the cotangents bars = (vors, divs, height)
        after it.  With stages (from swe_back_rk4_step_record) the forward is not recomputed.
        Returns (vors_bar, divs_bar, height_bar)."""
        out = [self.zeros_part() for _ in range(3)] if out is None else list(_seq(out, 3, "out"))
        a = [self._p(t, what="swe_back_rk4_step_adjoint input") for t in (x, y, z, area, vors, divs, height)]
        nb = [self._p(t, what="bars") for t in _seq(bars, 3, "bars")]
        ob = [self._p(t, what="out") for t in out]
        if stages is None:
            self._check(self._lib.sbslpm_swe_back_rk4_step_adjoint(self.h, *a, *nb, *ob), "swe_back_rk4_step_adjoint")
        else:
            self._check(self._lib.sbslpm_swe_back_rk4_step_adjoint_stages(self.h, *a, self._p(stages, 21, "stages"),
                                                                          *nb, *ob),
                        "swe_back_rk4_step_adjoint_stages")
        return out[0], out[1], out[2]

    def swe_find_departure_points(self, x, y, z, vels, dt, dep=None):
        """swe_find_departure_points, swe_time_step.cpp:422-556."""
        dep = self.zeros_part(3) if dep is None else dep
        _keep, arr = self._stages(vels, "vels")
        self._check(self._lib.sbslpm_swe_find_departure_points(
            self.h, *[self._p(t, what="x, y, z") for t in (x, y, z)], arr, float(dt), *self._p3(dep, "dep")),
            "swe_departure_points")
        return dep

    def swe_departure_to_target(self, dep, vor, div, h, tracers=None):
        """swe_departure_to_target, swe_time_step.cpp:19-74; returns (new_vor, new_div, new_h, new_tracers)."""
        out = [self.zeros_part() for _ in range(3)]
        ntr = 0 if tracers is None else tracers.shape[-1]
        new_tracers = self.torch.zeros_like(tracers) if ntr else None
        tr = [self._p(tracers, ntr, "tracers"), self._p(new_tracers, ntr, "new_tracers")] if ntr else [None, None]
        self._check(self._lib.sbslpm_swe_departure_to_target(
            self.h, *self._p3(dep, "dep"), *[self._p(t, what="vor, div, h") for t in (vor, div, h)], tr[0], ntr,
            *self._p3(out, "out"), tr[1]), "swe_departure_to_target")
        return out[0], out[1], out[2], new_tracers

    def bve_back_rk4_step(self, x, y, z, area, vors, vel, tracers=None, time=0.0, proxy_source_vors=None):
        """bve_back_rk4_step, bve_time_step.cpp:311-419: vors/tracers updated in place, vel <- stage-0 velocity."""
        ntr = self.cfg.tracer_count
        tr = self._p(tracers, ntr, "tracers") if ntr > 0 else None
        psv = None if proxy_source_vors is None else self._x(proxy_source_vors, "proxy_source_vors")
        a = [self._p(t, what="bve_back_rk4_step input") for t in (x, y, z, area, vors)]
        self._check(self._lib.sbslpm_bve_back_rk4_step(self.h, *a, psv, *self._p3(vel, "vel"), tr, float(time)),
                    "bve_back_rk4_step")
        return vors


class IrregContext:
    """Device handle of an irregular particle tree (the lat-lon path's _ll operators)."""

    def __init__(self, ctx, x, y, z, leaf_size):
        self.ctx, self._lib, self.torch = ctx, ctx._lib, ctx.torch
        a = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (x, y, z)]
        h = ctypes.c_void_p()
        ctx._check(self._lib.sbslpm_irreg_create(ctx.h, a[0].size, *[_dp(v) for v in a], int(leaf_size),
                                                 int(ctx.cfg.interp_degree), ctypes.byref(h)), "irreg_create")
        self.h = h
        s = np.zeros(6, dtype=np.int64)
        self._lib.sbslpm_irreg_sizes(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        self.P, self.levels, self.n, self.N, self.I, self.tiles = [int(v) for v in s]

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.sbslpm_irreg_destroy(self.h)
            self.h = None

    def _zeros(self, *shape):
        return self.torch.zeros(shape, dtype=self.torch.float64, device="cuda")

    def _pt(self, t, what="point array"):
        return _ptr(t, self.N, what)

    def _px(self, t, what="proxy array"):
        return _ptr(t, self.P * self.n, what)

    def upward_pass(self, area, pots, proxy=None):
        """upward_pass_ll: proxy charges [P, n] (assigned)."""
        proxy = self._zeros(self.P, self.n) if proxy is None else proxy
        self.ctx._check(self._lib.sbslpm_irreg_upward_pass(self.h, self._pt(area, "area"), self._pt(pots, "pots"),
                                                           self._px(proxy, "proxy")), "irreg_upward_pass")
        return proxy

    def downward_pass_3(self, proxy, sols=None):
        """downward_pass_3_ll: proxy (3 tensors [P, n]) pushed down in place; sols += L2P."""
        sols = [self._zeros(self.N) for _ in range(3)] if sols is None else sols
        self.ctx._check(self._lib.sbslpm_irreg_downward_pass_3(
            self.h, *[self._px(p, "proxy") for p in _seq(proxy, 3, "proxy")],
            *[self._pt(t, "sols") for t in _seq(sols, 3, "sols")]), "irreg_downward_pass_3")
        return sols

    def swe_vel_interactions(self, proxy_vors, proxy_divs, vors, divs, area, target_pots=None, vel=None):
        """swe_vel_interactions_ll: target_pots (3 x [P, n]) and vel (3 x [N]) accumulated."""
        target_pots = [self._zeros(self.P, self.n) for _ in range(3)] if target_pots is None else target_pots
        vel = [self._zeros(self.N) for _ in range(3)] if vel is None else vel
        self.ctx._check(self._lib.sbslpm_irreg_swe_vel_interactions(
            self.h, *[self._px(t, "target_pots") for t in _seq(target_pots, 3, "target_pots")],
            *[self._pt(t, "vel") for t in _seq(vel, 3, "vel")], self._px(proxy_vors, "proxy_vors"),
            self._px(proxy_divs, "proxy_divs"), self._pt(vors, "vors"), self._pt(divs, "divs"),
            self._pt(area, "area")), "irreg_swe_vel_interactions")
        return target_pots, vel

    def swe_velocity(self, area, vors, divs, vel=None):
        vel = [self._zeros(self.N) for _ in range(3)] if vel is None else vel
        self.ctx._check(self._lib.sbslpm_irreg_swe_velocity(
            self.h, self._pt(area, "area"), self._pt(vors, "vors"), self._pt(divs, "divs"),
            *[self._pt(t, "vel") for t in _seq(vel, 3, "vel")]), "irreg_swe_velocity")
        return vel

    def swe_vel_tend_ll(self, grid, x, y, z, area, vors, divs, height, topo, omega=None, vel=None, tend=None):
        """swe_vel_tend_computation on this tree (built from the grid's points): returns
        (vel [3], (vor_tend, div_tend, height_tend))."""
        from .latlon import OMEGA
        vel = [self._zeros(self.N) for _ in range(3)] if vel is None else _seq(vel, 3, "vel")
        tend = [self._zeros(self.N) for _ in range(3)] if tend is None else _seq(tend, 3, "tend")
        nlat, nlon, dx = int(grid[0]), int(grid[1]), float(grid[2])
        a = [self._pt(t, "swe_vel_tend_ll array") for t in (x, y, z, area, vors, divs, height, topo, *vel, *tend)]
        self.ctx._check(self._lib.sbslpm_irreg_swe_vel_tend_ll(
            self.h, nlat, nlon, dx, float(OMEGA if omega is None else omega), *a), "irreg_swe_vel_tend_ll")
        return vel, tend

    def swe_back_rk4_step_ll(self, grid, delta_t, x, y, z, area, vors, divs, height, topo, omega=None, vel=None,
                             time=0.0):
        """swe_back_rk4_step_ll: one step of (vors, divs, height) in place; returns vel, the
        start state's velocity."""
        from .latlon import OMEGA
        vel = [self._zeros(self.N) for _ in range(3)] if vel is None else _seq(vel, 3, "vel")
        nlat, nlon, dx = int(grid[0]), int(grid[1]), float(grid[2])
        a = [self._pt(t, "swe_back_rk4_step_ll array") for t in (x, y, z, area, vors, divs, height, topo, *vel)]
        self.ctx._check(self._lib.sbslpm_irreg_swe_back_rk4_step_ll(
            self.h, nlat, nlon, dx, float(delta_t), float(OMEGA if omega is None else omega), *a, float(time)),
            "irreg_swe_back_rk4_step_ll")
        return vel

    def bve_velocity(self, area, vor, vel=None):
        vel = [self._zeros(self.N) for _ in range(3)] if vel is None else vel
        self.ctx._check(self._lib.sbslpm_irreg_bve_velocity(
            self.h, self._pt(area, "area"), self._pt(vor, "vor"), *[self._pt(t, "vel") for t in _seq(vel, 3, "vel")]),
            "irreg_bve_velocity")
        return vel
