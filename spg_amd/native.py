"""ctypes bindings of the two native libraries.

  libspg_vbd.so  — include/spg_vbd.h, spg_vbd_learned.h, spg_xpbd.h: the GPU VBD and XPBD solvers'
                   C-ABI (the drop-in boundary)
  libspg_host.so — include/spg_host.h: the C++ mirror of spg's ParticleGroup / energies / BaseSolver

There is no CPU or PyTorch fallback: if the libraries are not built, every entry point raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")
VBD_LIB = os.path.join(LIBDIR, "libspg_vbd.so")
HOST_LIB = os.path.join(LIBDIR, "libspg_host.so")

SPRING, MEMBRANE_STVK, DISCRETE_BENDING, STABLE_NEO_HOOKEAN, SPRING_ANCHOR = 1, 2, 3, 4, 5
(SPRING_CONTINUUM, SPRING_SQUARED_CONTINUUM, MEMBRANE_BARAFF_WITKIN, MEMBRANE_CHOI, QUADRATIC_BENDING,
 BARAFF_WITKIN_BENDING, STVK, STABLE_SQUARED_NEO_HOOKEAN) = range(6, 14)
KIND = {"spring": SPRING, "stvk": MEMBRANE_STVK, "bending": DISCRETE_BENDING, "snh": STABLE_NEO_HOOKEAN,
        "anchor": SPRING_ANCHOR, "spring_continuum": SPRING_CONTINUUM,
        "spring_squared_continuum": SPRING_SQUARED_CONTINUUM, "membrane_bw": MEMBRANE_BARAFF_WITKIN,
        "membrane_choi": MEMBRANE_CHOI, "bending_quadratic": QUADRATIC_BENDING, "bending_bw": BARAFF_WITKIN_BENDING,
        "stvk_tet": STVK, "ssnh": STABLE_SQUARED_NEO_HOOKEAN}
STENCIL_SIZE = {SPRING: 2, MEMBRANE_STVK: 3, DISCRETE_BENDING: 4, STABLE_NEO_HOOKEAN: 4, SPRING_ANCHOR: 1,
                SPRING_CONTINUUM: 2, SPRING_SQUARED_CONTINUUM: 2, MEMBRANE_BARAFF_WITKIN: 3, MEMBRANE_CHOI: 3,
                QUADRATIC_BENDING: 4, BARAFF_WITKIN_BENDING: 4, STVK: 4, STABLE_SQUARED_NEO_HOOKEAN: 4}
PARAMS = {SPRING: 2, MEMBRANE_STVK: 13, DISCRETE_BENDING: 2, STABLE_NEO_HOOKEAN: 12, SPRING_ANCHOR: 4,
          SPRING_CONTINUUM: 2, SPRING_SQUARED_CONTINUUM: 2, MEMBRANE_BARAFF_WITKIN: 7, MEMBRANE_CHOI: 8,
          QUADRATIC_BENDING: 5, BARAFF_WITKIN_BENDING: 2, STVK: 11, STABLE_SQUARED_NEO_HOOKEAN: 11}
XPBD_PARAMS = {SPRING: 2, MEMBRANE_STVK: 13, DISCRETE_BENDING: 2, STABLE_NEO_HOOKEAN: 14, SPRING_ANCHOR: 4,
               SPRING_CONTINUUM: 2, SPRING_SQUARED_CONTINUUM: 2, MEMBRANE_BARAFF_WITKIN: 7, MEMBRANE_CHOI: 8,
               QUADRATIC_BENDING: 5, BARAFF_WITKIN_BENDING: 2, STVK: 11, STABLE_SQUARED_NEO_HOOKEAN: 11}
ERR_NONFINITE = 5
ERR_STATE = 4

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_fp = ctypes.POINTER(ctypes.c_float)


class VbdEnergy(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("n_stencils", ctypes.c_int64), ("stencils", _ip), ("params", _dp)]


class VbdObject(ctypes.Structure):
    _fields_ = [("n_particles", ctypes.c_int64), ("x", _dp), ("v", _dp), ("mass", _dp),
                ("n_energies", ctypes.c_int32), ("energies", ctypes.POINTER(VbdEnergy)), ("x0", _dp)]


class VbdLaunch(ctypes.Structure):
    _fields_ = [("lanes_per_vertex", ctypes.c_int32), ("block_size", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("lds_staging", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class VbdStats(ctypes.Structure):
    _fields_ = [("solve_launches", ctypes.c_int64), ("solve_ms", ctypes.c_double), ("vertex_updates", ctypes.c_int64),
                ("algorithmic_bytes", ctypes.c_double), ("fp64_fallbacks", ctypes.c_int64),
                ("skipped_blocks", ctypes.c_int64), ("prediction_fallbacks", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class XpbdEnergy(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("n_stencils", ctypes.c_int64), ("stencils", _ip), ("params", _dp)]


class XpbdObject(ctypes.Structure):
    _fields_ = [("n_particles", ctypes.c_int64), ("x", _dp), ("v", _dp), ("inv_mass", _dp),
                ("n_energies", ctypes.c_int32), ("energies", ctypes.POINTER(XpbdEnergy))]


class XpbdLaunch(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int32), ("use_graph", ctypes.c_int32), ("group_runs", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class XpbdStats(ctypes.Structure):
    _fields_ = [("projection_launches", ctypes.c_int64), ("projection_ms", ctypes.c_double),
                ("stencil_projections", ctypes.c_int64), ("algorithmic_bytes", ctypes.c_double),
                ("fp64_solves", ctypes.c_int64), ("type_launches", ctypes.c_int64 * 14),
                ("type_ms", ctypes.c_double * 14), ("type_bytes", ctypes.c_double * 14)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_[:5]}
        names = {t: k for k, t in KIND.items()}
        names[0] = "small_group_runs"
        d["per_type"] = {names[t]: {"launches": self.type_launches[t], "ms": self.type_ms[t],
                                    "bytes": self.type_bytes[t]} for t in names if self.type_launches[t]}
        return d


class ImplicitStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("linear_solves", ctypes.c_int64), ("cg_iterations", ctypes.c_int64),
                ("newton_iterations", ctypes.c_int64), ("failed_substeps", ctypes.c_int64),
                ("current_dt", ctypes.c_double), ("last_cg_error", ctypes.c_double), ("assembly_ms", ctypes.c_double),
                ("cg_ms", ctypes.c_double), ("spmv_launches", ctypes.c_int64), ("spmv_bytes", ctypes.c_double),
                ("spmv_sampled_ms", ctypes.c_double), ("spmv_sampled", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RigidEnergy(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("n_stencils", ctypes.c_int64), ("stencils", _ip), ("params", _dp)]


class RigidGroup(ctypes.Structure):
    _fields_ = [("n_bodies", ctypes.c_int64), ("x", _dp), ("theta", _dp), ("v", _dp), ("omega", _dp),
                ("mass", _dp), ("inertia", _dp), ("n_energies", ctypes.c_int32),
                ("energies", ctypes.POINTER(RigidEnergy))]


class RigidStats(ctypes.Structure):
    _fields_ = [("n_colors", ctypes.c_int32), ("n_bodies", ctypes.c_int64), ("n_elements", ctypes.c_int64),
                ("incidence_entries", ctypes.c_int64), ("skipped_blocks", ctypes.c_int64),
                ("solve_bytes", ctypes.c_int64)]


RIGID_KIND = {"rb_spring": 21, "rb_spring_squared": 22, "rb_anchor": 23, "rb_alignment": 24}
RIGID_STENCIL = {21: 2, 22: 2, 23: 1, 24: 2}
RIGID_PARAMS = {21: 8, 22: 8, 23: 7, 24: 1}


class NativeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_vbd = None
_host = None


def _proto(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def vbd_lib():
    global _vbd
    if _vbd is None:
        if not os.path.exists(VBD_LIB):
            raise RuntimeError(f"{VBD_LIB} is not built: run `python -m spg_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(VBD_LIB)
        c = ctypes.c_void_p
        i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        _proto(L, "spg_vbd_create", ctypes.c_int, [ctypes.POINTER(c), ctypes.POINTER(VbdObject), i32, i32])
        _proto(L, "spg_vbd_destroy", None, [c])
        _proto(L, "spg_vbd_last_error", ctypes.c_char_p, [c])
        _proto(L, "spg_vbd_num_objects", ctypes.c_int, [c, _ip])
        _proto(L, "spg_vbd_num_particles", ctypes.c_int, [c, i32, _lp])
        _proto(L, "spg_vbd_num_colors", ctypes.c_int, [c, i32, _ip])
        _proto(L, "spg_vbd_get_colors", ctypes.c_int, [c, i32, _ip])
        _proto(L, "spg_vbd_set_state", ctypes.c_int, [c, i32, _dp, _dp])
        _proto(L, "spg_vbd_get_state", ctypes.c_int, [c, i32, _dp, _dp])
        _proto(L, "spg_vbd_step", ctypes.c_int, [c, dbl, i32, i32, _dp])
        _proto(L, "spg_vbd_begin_substep", ctypes.c_int, [c, dbl, _dp])
        _proto(L, "spg_vbd_set_guess", ctypes.c_int, [c, i32, _dp])
        _proto(L, "spg_vbd_iterate", ctypes.c_int, [c, i32])
        _proto(L, "spg_vbd_end_substep", ctypes.c_int, [c])
        _proto(L, "spg_vbd_save_state", ctypes.c_int, [c])
        _proto(L, "spg_vbd_restore_state", ctypes.c_int, [c])
        _proto(L, "spg_vbd_residual", ctypes.c_int, [c, _dp])
        _proto(L, "spg_vbd_bench_barrier", ctypes.c_int, [i32, i32, i32, i32, i32, _dp, _dp])
        _proto(L, "spg_vbd_residual_noalloc", ctypes.c_int, [c, _dp])
        _proto(L, "spg_vbd_residual_at_inertial", ctypes.c_int, [c, _dp])
        _proto(L, "spg_vbd_iterate_to_tolerance", ctypes.c_int, [c, dbl, i32, i32, _ip, _dp])
        _proto(L, "spg_vbd_get_launch", ctypes.c_int, [c, ctypes.POINTER(VbdLaunch)])
        _proto(L, "spg_vbd_set_launch", ctypes.c_int, [c, ctypes.POINTER(VbdLaunch)])
        _proto(L, "spg_vbd_autotune", ctypes.c_int, [c, i32, ctypes.POINTER(VbdLaunch)])
        _proto(L, "spg_vbd_get_stats", ctypes.c_int, [c, ctypes.POINTER(VbdStats)])
        _proto(L, "spg_vbd_reset_stats", ctypes.c_int, [c])
        _proto(L, "spg_vbd_set_timing", ctypes.c_int, [c, i32])
        _proto(L, "spg_vbd_synchronize", ctypes.c_int, [c])
        _proto(L, "spg_vbd_color_vertices", ctypes.c_int,
               [i64, i32, ctypes.POINTER(_ip), _lp, _ip, _ip])
        _proto(L, "spg_vbd_element_derivatives", ctypes.c_int, [i32, i64, _dp, _dp, i32, i32, i32, _dp, _dp])
        _proto(L, "spg_vbd_bench_copy", ctypes.c_int, [i64, i32, i32, _dp])
        _proto(L, "spg_vbd_bench_residual", ctypes.c_int, [c, i32, _dp, _dp])
        # learned warm start (include/spg_vbd_learned.h)
        u64 = ctypes.c_uint64
        _proto(L, "spg_mlp_create", ctypes.c_int, [ctypes.POINTER(c), i32, u64])
        _proto(L, "spg_mlp_destroy", None, [c])
        _proto(L, "spg_mlp_last_error", ctypes.c_char_p, [c])
        _proto(L, "spg_mlp_get_params", ctypes.c_int, [c, _fp])
        _proto(L, "spg_mlp_set_params", ctypes.c_int, [c, _fp])
        _proto(L, "spg_mlp_get_normalization", ctypes.c_int, [c, _fp, _fp])
        _proto(L, "spg_mlp_set_normalization", ctypes.c_int, [c, _fp, _fp])
        _proto(L, "spg_mlp_set_history_base", ctypes.c_int, [c, i32])
        _proto(L, "spg_mlp_get_history_base", ctypes.c_int, [c, _ip])
        _proto(L, "spg_mlp_forward", ctypes.c_int, [c, _fp, i64, _fp])
        _proto(L, "spg_mlp_loss_grad", ctypes.c_int, [c, _fp, _fp, i64, _dp, _fp])
        _proto(L, "spg_mlp_train", ctypes.c_int, [c, _fp, _fp, i64, i32, i32, ctypes.c_float, u64, i32, _dp])
        _proto(L, "spg_vbd_features", ctypes.c_int, [c, i32, _fp])
        _proto(L, "spg_vbd_warm_start", ctypes.c_int, [c, c, i32, i32, _ip, _dp])
        _proto(L, "spg_vbd_history_start", ctypes.c_int, [c, i32, _ip, _dp])
        _proto(L, "spg_vbd_start_rho", ctypes.c_int, [c, _dp])
        # XPBD (include/spg_xpbd.h)
        _proto(L, "spg_xpbd_create", ctypes.c_int, [ctypes.POINTER(c), ctypes.POINTER(XpbdObject), i32, i32, i32])
        _proto(L, "spg_xpbd_destroy", None, [c])
        _proto(L, "spg_xpbd_last_error", ctypes.c_char_p, [c])
        _proto(L, "spg_xpbd_num_objects", ctypes.c_int, [c, _ip])
        _proto(L, "spg_xpbd_num_particles", ctypes.c_int, [c, i32, _lp])
        _proto(L, "spg_xpbd_num_groups", ctypes.c_int, [c, i32, i32, _ip])
        _proto(L, "spg_xpbd_get_groups", ctypes.c_int, [c, i32, i32, _ip])
        _proto(L, "spg_xpbd_num_launches", ctypes.c_int, [c, _ip])
        _proto(L, "spg_xpbd_set_state", ctypes.c_int, [c, i32, _dp, _dp])
        _proto(L, "spg_xpbd_get_state", ctypes.c_int, [c, i32, _dp, _dp])
        _proto(L, "spg_xpbd_step", ctypes.c_int, [c, dbl, i32, _dp])
        _proto(L, "spg_xpbd_save_state", ctypes.c_int, [c])
        _proto(L, "spg_xpbd_restore_state", ctypes.c_int, [c])
        _proto(L, "spg_xpbd_get_launch", ctypes.c_int, [c, ctypes.POINTER(XpbdLaunch)])
        _proto(L, "spg_xpbd_set_launch", ctypes.c_int, [c, ctypes.POINTER(XpbdLaunch)])
        _proto(L, "spg_xpbd_get_stats", ctypes.c_int, [c, ctypes.POINTER(XpbdStats)])
        _proto(L, "spg_xpbd_reset_stats", ctypes.c_int, [c])
        _proto(L, "spg_xpbd_set_timing", ctypes.c_int, [c, i32])
        _proto(L, "spg_xpbd_synchronize", ctypes.c_int, [c])
        _proto(L, "spg_xpbd_color_stencils", ctypes.c_int, [i64, i32, _ip, _ip])
        _proto(L, "spg_xpbd_project_stencils", ctypes.c_int, [i32, i64, _dp, _dp, _dp, dbl, i32, i32, _dp])
        # implicit Euler family (include/spg_implicit.h)
        _proto(L, "spg_implicit_create", ctypes.c_int, [ctypes.POINTER(c), ctypes.POINTER(VbdObject), i32, i32, i32])
        _proto(L, "spg_implicit_destroy", None, [c])
        _proto(L, "spg_implicit_last_error", ctypes.c_char_p, [c])
        _proto(L, "spg_implicit_num_dofs", ctypes.c_int, [c, _lp])
        _proto(L, "spg_implicit_num_blocks", ctypes.c_int, [c, _lp])
        _proto(L, "spg_implicit_set_state", ctypes.c_int, [c, i32, _dp, _dp])
        _proto(L, "spg_implicit_get_state", ctypes.c_int, [c, i32, _dp, _dp])
        _proto(L, "spg_implicit_step", ctypes.c_int, [c, dbl, i32, _dp])
        _proto(L, "spg_implicit_reset", ctypes.c_int, [c])
        _proto(L, "spg_implicit_system_force", ctypes.c_int, [c, _dp, _dp])
        _proto(L, "spg_implicit_stiffness", ctypes.c_int, [c, _lp, _ip, _dp])
        _proto(L, "spg_implicit_cg_solve", ctypes.c_int, [c, dbl, dbl, _dp, dbl, i32, _dp, _ip, _dp])
        _proto(L, "spg_implicit_bench_spmv", ctypes.c_int, [c, dbl, dbl, i32, _dp])
        _proto(L, "spg_implicit_element_derivatives", ctypes.c_int, [i32, ctypes.c_int64, _dp, _dp, _dp, _dp, i32])
        _proto(L, "spg_implicit_get_stats", ctypes.c_int, [c, ctypes.POINTER(ImplicitStats)])
        _proto(L, "spg_implicit_reset_stats", ctypes.c_int, [c])
        _proto(L, "spg_implicit_set_timing", ctypes.c_int, [c, i32])
        _proto(L, "spg_implicit_synchronize", ctypes.c_int, [c])
        # rigid bodies (include/spg_rigid.h)
        _proto(L, "spg_rigid_create", ctypes.c_int, [ctypes.POINTER(c), ctypes.POINTER(RigidGroup), i32, i32])
        _proto(L, "spg_rigid_destroy", None, [c])
        _proto(L, "spg_rigid_last_error", ctypes.c_char_p, [c])
        _proto(L, "spg_rigid_step", ctypes.c_int, [c, dbl, i32, i32, _dp])
        _proto(L, "spg_rigid_get_state", ctypes.c_int, [c, i32, _dp, _dp, _dp, _dp])
        _proto(L, "spg_rigid_set_state", ctypes.c_int, [c, i32, _dp, _dp, _dp, _dp])
        _proto(L, "spg_rigid_get_colors", ctypes.c_int, [c, i32, _ip, _ip])
        _proto(L, "spg_rigid_get_rotations", ctypes.c_int, [c, i32, _dp])
        _proto(L, "spg_rigid_set_rotations", ctypes.c_int, [c, i32, _dp])
        _proto(L, "spg_rigid_get_stats", ctypes.c_int, [c, ctypes.POINTER(RigidStats)])
        _proto(L, "spg_rigid_set_graph", ctypes.c_int, [c, i32])
        _proto(L, "spg_rigid_profile_substep", ctypes.c_int, [c, dbl, i32, _dp, _dp, _ip])
        _proto(L, "spg_rigid_element_derivatives", ctypes.c_int, [i32, i64, _dp, _dp, _dp, _dp, _dp, i32])
        _vbd = L
    return _vbd


def host_lib():
    global _host
    if _host is None:
        vbd_lib()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built: run `python -m spg_amd.build`")
        L = ctypes.CDLL(HOST_LIB)
        c = ctypes.c_void_p
        i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        _proto(L, "spg_host_last_error", ctypes.c_char_p, [])
        _proto(L, "spg_pg_create", c, [])
        _proto(L, "spg_pg_destroy", None, [c])
        _proto(L, "spg_pg_add_particles", ctypes.c_int, [c, i64, _dp, _dp, _dp])
        _proto(L, "spg_pg_add_energy", ctypes.c_int, [c, i32, i64, _ip, _dp, i32, _ip])
        _proto(L, "spg_pg_prepare_energy", ctypes.c_int, [c, i32])
        _proto(L, "spg_pg_num_particles", ctypes.c_int, [c, _lp])
        _proto(L, "spg_pg_energy_params", ctypes.c_int, [c, i32, _dp])
        _proto(L, "spg_pg_energy_xpbd_params", ctypes.c_int, [c, i32, _dp])
        _proto(L, "spg_pg_inv_masses", ctypes.c_int, [c, _dp])
        _proto(L, "spg_pg_get_state", ctypes.c_int, [c, _dp, _dp])
        _proto(L, "spg_pg_set_state", ctypes.c_int, [c, _dp, _dp])
        _proto(L, "spg_trimesh_bending_stencils", ctypes.c_int, [i64, i64, _ip, _ip, _lp])
        _proto(L, "spg_io_load_obj", ctypes.c_int, [ctypes.c_char_p, _lp, _lp, _dp, _ip])
        _proto(L, "spg_io_load_msh", ctypes.c_int, [ctypes.c_char_p, _lp, _lp, _dp, _ip])
        _proto(L, "spg_solver_create_vbd_gpu", c, [i32])
        _proto(L, "spg_solver_create_xpbd_gpu", c, [i32, i32])
        _proto(L, "spg_solver_xpbd_context", c, [c])
        _proto(L, "spg_solver_create_implicit_gpu", c, [i32, i32])
        _proto(L, "spg_solver_implicit_context", c, [c])
        _proto(L, "spg_solver_destroy", None, [c])
        _proto(L, "spg_solver_add_object", ctypes.c_int, [c, c])
        _proto(L, "spg_rb_create", c, [])
        _proto(L, "spg_rb_destroy", None, [c])
        _proto(L, "spg_rb_add_bodies", ctypes.c_int, [c, i64, _dp, _dp, _dp, _dp, _dp, _dp])
        _proto(L, "spg_rb_add_energy", ctypes.c_int, [c, i32, i64, _ip, _dp, _ip])
        _proto(L, "spg_rb_num_bodies", ctypes.c_int, [c, _lp])
        _proto(L, "spg_rb_get_state", ctypes.c_int, [c, _dp, _dp, _dp, _dp, _dp])
        _proto(L, "spg_rb_set_state", ctypes.c_int, [c, _dp, _dp, _dp, _dp])
        _proto(L, "spg_solver_add_rb_object", ctypes.c_int, [c, c])
        _proto(L, "spg_solver_get_rb_state", ctypes.c_int, [c, i32, _dp, _dp, _dp, _dp, _dp])
        _proto(L, "spg_solver_rigid_context", c, [c])
        _proto(L, "spg_solver_set_dt", ctypes.c_int, [c, dbl])
        _proto(L, "spg_solver_set_num_substeps", ctypes.c_int, [c, i32])
        _proto(L, "spg_solver_set_gravity", ctypes.c_int, [c, _dp])
        _proto(L, "spg_solver_set_iterations", ctypes.c_int, [c, i32])
        _proto(L, "spg_solver_set_resident_state", ctypes.c_int, [c, i32])
        _proto(L, "spg_solver_step", ctypes.c_int, [c])
        _proto(L, "spg_solver_reset", ctypes.c_int, [c])
        _proto(L, "spg_solver_require_precomputation_update", ctypes.c_int, [c])
        _proto(L, "spg_solver_sync_to_host", ctypes.c_int, [c])
        _proto(L, "spg_solver_get_state", ctypes.c_int, [c, i32, _dp, _dp])
        _proto(L, "spg_solver_context", c, [c])
        _host = L
    return _host


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _fl(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(shape) if shape is not None else a


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def _i32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a.reshape(shape) if shape is not None else a


def _kind(k):
    return KIND[k] if isinstance(k, str) else int(k)


# ------------------------------------------------------------------------------------------------
# Host mirror: ParticleGroup / TriangleMesh / VBDGpu
# ------------------------------------------------------------------------------------------------

def _host_check(rc):
    if rc != 0:
        raise NativeError(rc, host_lib().spg_host_last_error().decode())


def trimesh_bending_stencils(n_vertices, faces):
    """spg::TriangleMesh(faces).bendingStencils(): {e0, e1, opp0, opp1} for interior edges."""
    L = host_lib()
    faces = _i32(faces, (-1, 3))
    n = ctypes.c_int64(0)
    _host_check(L.spg_trimesh_bending_stencils(n_vertices, faces.shape[0], _i(faces), None, ctypes.byref(n)))
    out = np.zeros((n.value, 4), dtype=np.int32)
    _host_check(L.spg_trimesh_bending_stencils(n_vertices, faces.shape[0], _i(faces), _i(out), ctypes.byref(n)))
    return out


class ParticleGroup:
    """spg::ParticleGroup built through the C++ mirror (addParticle, addStencil, preparePrecomputations)."""

    def __init__(self):
        self._L = host_lib()
        self.h = self._L.spg_pg_create()
        self.energy_types = []

    def __del__(self):
        if getattr(self, "h", None):
            self._L.spg_pg_destroy(self.h)
            self.h = None

    def add_particles(self, x, x0, m):
        x, x0, m = _f64(x, (-1, 3)), _f64(x0, (-1, 3)), _f64(m)
        _host_check(self._L.spg_pg_add_particles(self.h, x.shape[0], _d(x), _d(x0), _d(m)))

    def add_energy(self, kind, stencils, args, prepare=True):
        t = _kind(kind)
        st = _i32(stencils, (-1, STENCIL_SIZE[t]))
        a = _f64(args, (st.shape[0], -1))
        idx = ctypes.c_int32(0)
        _host_check(self._L.spg_pg_add_energy(self.h, t, st.shape[0], _i(st), _d(a), int(prepare), ctypes.byref(idx)))
        self.energy_types.append((t, st.shape[0], st))
        return idx.value

    def prepare_energy(self, index):
        _host_check(self._L.spg_pg_prepare_energy(self.h, index))

    def num_particles(self):
        n = ctypes.c_int64(0)
        _host_check(self._L.spg_pg_num_particles(self.h, ctypes.byref(n)))
        return n.value

    def energy_params(self, index):
        t, n, _ = self.energy_types[index]
        out = np.zeros((n, PARAMS[t]))
        _host_check(self._L.spg_pg_energy_params(self.h, index, _d(out)))
        return out

    def state(self):
        n = self.num_particles()
        x, v = np.zeros((n, 3)), np.zeros((n, 3))
        _host_check(self._L.spg_pg_get_state(self.h, _d(x), _d(v)))
        return x, v

    def xpbd_params(self, index):
        """Energy `index` in the compliance layout of spg_xpbd_energy (spg_xpbd.h)."""
        t, n, _ = self.energy_types[index]
        out = np.zeros((n, XPBD_PARAMS[t]))
        _host_check(self._L.spg_pg_energy_xpbd_params(self.h, index, _d(out)))
        return out

    def inv_masses(self):
        w = np.zeros(self.num_particles())
        _host_check(self._L.spg_pg_inv_masses(self.h, _d(w)))
        return w

    def set_state(self, x=None, v=None):
        _host_check(self._L.spg_pg_set_state(self.h, _d(None if x is None else _f64(x, (-1, 3))),
                                             _d(None if v is None else _f64(v, (-1, 3)))))

    def vbd_object(self, x=None, v=None, m=None):
        """Prepared arrays of this group in the layout of spg_vbd_object (for direct C-ABI use)."""
        xx, vv = self.state()
        return {"x": xx if x is None else x, "v": vv if v is None else v, "m": m,
                "energies": [(t, st, self.energy_params(i)) for i, (t, n, st) in enumerate(self.energy_types)]}


class RigidBodyGroup:
    """spg::RigidBodyGroup built through the C++ mirror (addBody, the rigid-body energies' addStencil)."""

    def __init__(self, scene=None):
        self._L = host_lib()
        self.h = self._L.spg_rb_create()
        if scene is not None:
            self.add_bodies(scene.x, scene.m, scene.theta, scene.inertia)
            for kind, st, args in scene.energies:
                self.add_energy(kind, st, args)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.spg_rb_destroy(self.h)
            self.h = None

    def add_bodies(self, x, m, theta, inertia, x0=None, theta0=None):
        x = _f64(x, (-1, 3))
        n = x.shape[0]
        m, th, inertia = _f64(m, (n,)), _f64(theta, (n, 3)), _f64(inertia, (n, 9))
        x0 = None if x0 is None else _f64(x0, (n, 3))
        th0 = None if theta0 is None else _f64(theta0, (n, 3))
        _host_check(self._L.spg_rb_add_bodies(self.h, n, _d(x), None if x0 is None else _d(x0), _d(m), _d(th),
                                              None if th0 is None else _d(th0), _d(inertia)))

    def add_energy(self, kind, stencils, args):
        t = RIGID_KIND.get(kind, kind)
        st = _i32(stencils, (-1, RIGID_STENCIL[t]))
        a = _f64(args, (st.shape[0], RIGID_PARAMS[t]))
        idx = ctypes.c_int32(0)
        _host_check(self._L.spg_rb_add_energy(self.h, t, st.shape[0], _i(st), _d(a), ctypes.byref(idx)))
        return idx.value

    def num_bodies(self):
        n = ctypes.c_int64(0)
        _host_check(self._L.spg_rb_num_bodies(self.h, ctypes.byref(n)))
        return n.value

    def state(self):
        """(x, theta, v, omega, rotation matrices)."""
        n = self.num_bodies()
        out = [np.zeros((n, 3)) for _ in range(4)] + [np.zeros((n, 3, 3))]
        _host_check(self._L.spg_rb_get_state(self.h, *[_d(a) for a in out]))
        return tuple(out)

    def set_state(self, x=None, theta=None, v=None, omega=None):
        n = self.num_bodies()
        arrs = [None if a is None else _f64(a, (n, 3)) for a in (x, theta, v, omega)]
        _host_check(self._L.spg_rb_set_state(self.h, *[None if a is None else _d(a) for a in arrs]))


class VBDGpu:
    """spg::solver::VBDGpu (the GPU drop-in for spg::solver::VBD) through the C++ mirror."""

    def __init__(self, device=0):
        self._L = host_lib()
        self.h = self._L.spg_solver_create_vbd_gpu(device)
        self._groups = []
        self._rb_groups = []

    def add_rb_object(self, rb):
        _host_check(self._L.spg_solver_add_rb_object(self.h, rb.h))
        self._rb_groups.append(rb)

    def rb_state(self, obj=0):
        """rbGroups()[obj]: (x, theta, v, omega, rotation matrices)."""
        n = self._rb_groups[obj].num_bodies()
        out = [np.zeros((n, 3)) for _ in range(4)] + [np.zeros((n, 3, 3))]
        _host_check(self._L.spg_solver_get_rb_state(self.h, obj, *[_d(a) for a in out]))
        return tuple(out)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.spg_solver_destroy(self.h)
            self.h = None

    def add_object(self, pg):
        _host_check(self._L.spg_solver_add_object(self.h, pg.h))
        self._groups.append(pg)

    def set_dt(self, dt):
        _host_check(self._L.spg_solver_set_dt(self.h, dt))

    def set_num_substeps(self, n):
        _host_check(self._L.spg_solver_set_num_substeps(self.h, n))

    def set_gravity(self, g):
        g = _f64(g)
        _host_check(self._L.spg_solver_set_gravity(self.h, _d(g)))

    def set_iterations(self, n):
        _host_check(self._L.spg_solver_set_iterations(self.h, n))

    def set_resident_state(self, resident):
        _host_check(self._L.spg_solver_set_resident_state(self.h, int(resident)))

    def step(self):
        _host_check(self._L.spg_solver_step(self.h))

    def reset(self):
        _host_check(self._L.spg_solver_reset(self.h))

    def require_precomputation_update(self):
        _host_check(self._L.spg_solver_require_precomputation_update(self.h))

    def sync_to_host(self):
        _host_check(self._L.spg_solver_sync_to_host(self.h))

    def state(self, obj=0):
        n = self._groups[obj].num_particles()
        x, v = np.zeros((n, 3)), np.zeros((n, 3))
        _host_check(self._L.spg_solver_get_state(self.h, obj, _d(x), _d(v)))
        return x, v

    def context(self):
        """The C-ABI context behind this solver (valid after the first step), borrowed."""
        p = self._L.spg_solver_context(self.h)
        if not p:
            raise RuntimeError("no context before the first step")
        return VbdContext._borrow(p, self)


class XPBDGpu(VBDGpu):
    """spg::solver::XPBDGpu (the GPU drop-in for spg::solver::XPBD) through the C++ mirror."""

    def __init__(self, parallel_gauss_seidel=True, device=0):
        self._L = host_lib()
        self.h = self._L.spg_solver_create_xpbd_gpu(device, int(parallel_gauss_seidel))
        self._groups = []
        self._rb_groups = []

    def set_iterations(self, n):
        raise NotImplementedError("XPBD projects every stencil once per substep (no iteration count)")

    def context(self):
        """The C-ABI XPBD context behind this solver (valid after the first step), borrowed."""
        p = self._L.spg_solver_xpbd_context(self.h)
        if not p:
            raise RuntimeError("no context before the first step")
        return XpbdContext._borrow(p, self)


class ImplicitEulerGpu(VBDGpu):
    """spg::solver::ImplicitEulerGpu (GPU drop-in for the implicit Euler family) through the C++ mirror."""

    def __init__(self, solver, device=0):
        self._L = host_lib()
        kind = IMPLICIT_SOLVERS[solver] if isinstance(solver, str) else int(solver)
        self.h = self._L.spg_solver_create_implicit_gpu(device, kind)
        if not self.h:
            raise NativeError(1, self._L.spg_host_last_error().decode())
        self._groups = []
        self._rb_groups = []

    def set_iterations(self, n):
        raise NotImplementedError("the implicit solvers have no iteration count")


# ------------------------------------------------------------------------------------------------
# Direct C-ABI context
# ------------------------------------------------------------------------------------------------

class VbdContext:
    """A spg_vbd_ctx: one or more particle groups advanced together on one GPU."""

    def __init__(self, objects, device=0):
        """objects: list of dicts with x (n,3), v (n,3) or None, m (n,), energies [(type, stencils, params)]."""
        self._L = vbd_lib()
        keep = []
        objs = (VbdObject * len(objects))()
        for k, o in enumerate(objects):
            x = _f64(o["x"], (-1, 3))
            v = None if o.get("v") is None else _f64(o["v"], (-1, 3))
            x0 = None if o.get("x0") is None else _f64(o["x0"], (-1, 3))
            m = _f64(o["m"])
            ens = (VbdEnergy * max(1, len(o["energies"])))()
            for j, (t, st, prm) in enumerate(o["energies"]):
                t = _kind(t)
                st = _i32(st, (-1, STENCIL_SIZE[t]))
                prm = _f64(prm, (st.shape[0], PARAMS[t]))
                ens[j] = VbdEnergy(t, st.shape[0], _i(st), _d(prm))
                keep += [st, prm]
            objs[k] = VbdObject(x.shape[0], _d(x), _d(v), _d(m), len(o["energies"]), ens, _d(x0))
            keep += [x, v, x0, m, ens]
        h = ctypes.c_void_p()
        rc = self._L.spg_vbd_create(ctypes.byref(h), objs, len(objects), device)
        if rc != 0:
            raise NativeError(rc, self._L.spg_vbd_last_error(None).decode())
        self.h = h.value
        self._owned = True
        self._owner = None
        self.n_objects = len(objects)
        self._n = [int(_f64(o["x"]).size // 3) for o in objects]

    @classmethod
    def _borrow(cls, ptr, owner):
        self = cls.__new__(cls)
        self._L = vbd_lib()
        self.h = ptr
        self._owned = False
        self._owner = owner
        n = ctypes.c_int32(0)
        self._L.spg_vbd_num_objects(ptr, ctypes.byref(n))
        self.n_objects = n.value
        self._n = []
        for o in range(self.n_objects):
            k = ctypes.c_int64(0)
            self._L.spg_vbd_num_particles(ptr, o, ctypes.byref(k))
            self._n.append(k.value)
        return self

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "h", None):
            self._L.spg_vbd_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise NativeError(rc, self._L.spg_vbd_last_error(self.h).decode())

    def num_particles(self, obj=0):
        return self._n[obj]

    def colors(self, obj=0):
        out = np.zeros(self._n[obj], dtype=np.int32)
        self._check(self._L.spg_vbd_get_colors(self.h, obj, _i(out)))
        return out

    def num_colors(self, obj=0):
        n = ctypes.c_int32(0)
        self._check(self._L.spg_vbd_num_colors(self.h, obj, ctypes.byref(n)))
        return n.value

    def state(self, obj=0):
        x, v = np.zeros((self._n[obj], 3)), np.zeros((self._n[obj], 3))
        self._check(self._L.spg_vbd_get_state(self.h, obj, _d(x), _d(v)))
        return x, v

    def set_state(self, obj, x=None, v=None):
        self._check(self._L.spg_vbd_set_state(self.h, obj, _d(None if x is None else _f64(x, (-1, 3))),
                                              _d(None if v is None else _f64(v, (-1, 3)))))

    def step(self, dt, nsubsteps=1, iterations=1, gravity=(0.0, -9.8, 0.0)):
        g = _f64(gravity)
        self._check(self._L.spg_vbd_step(self.h, dt, nsubsteps, iterations, _d(g)))

    def begin_substep(self, h, gravity):
        g = _f64(gravity)
        self._check(self._L.spg_vbd_begin_substep(self.h, h, _d(g)))

    def set_guess(self, obj, x):
        self._check(self._L.spg_vbd_set_guess(self.h, obj, _d(_f64(x, (-1, 3)))))

    def iterate(self, iterations):
        self._check(self._L.spg_vbd_iterate(self.h, iterations))

    def end_substep(self):
        self._check(self._L.spg_vbd_end_substep(self.h))

    def save_state(self):
        self._check(self._L.spg_vbd_save_state(self.h))

    def restore_state(self):
        self._check(self._L.spg_vbd_restore_state(self.h))

    def residual(self):
        out = np.zeros(self.n_objects)
        self._check(self._L.spg_vbd_residual(self.h, _d(out)))
        return out

    def residual_noalloc(self):
        """spg_vbd_residual_noalloc: the same norms with the library's static work memory (no allocation)."""
        out = np.zeros(self.n_objects)
        self._check(self._L.spg_vbd_residual_noalloc(self.h, _d(out)))
        return out

    def residual_at_inertial(self):
        out = np.zeros(self.n_objects)
        self._check(self._L.spg_vbd_residual_at_inertial(self.h, _d(out)))
        return out

    def bench_residual(self, repeats=20):
        """spg_vbd_bench_residual: (mean us of one FP64 residual evaluation, its compulsory bytes)."""
        us, nbytes = ctypes.c_double(0), ctypes.c_double(0)
        self._check(self._L.spg_vbd_bench_residual(self.h, int(repeats), ctypes.byref(us), ctypes.byref(nbytes)))
        return us.value, nbytes.value

    def iterate_to_tolerance(self, rel_tol, max_iterations=1000, check_every=1):
        it = ctypes.c_int32(0)
        rho = ctypes.c_double(0.0)
        self._check(self._L.spg_vbd_iterate_to_tolerance(self.h, rel_tol, max_iterations, check_every,
                                                          ctypes.byref(it), ctypes.byref(rho)))
        return it.value, rho.value

    def features(self, obj=-1):
        """Learned-warm-start input rows of the open substep: (n, 24) float32 for object `obj`, or for
        all objects concatenated (obj = -1)."""
        n = sum(self._n) if obj < 0 else self._n[obj]
        out = np.zeros((n, MLP_FEATURES), dtype=np.float32)
        self._check(self._L.spg_vbd_features(self.h, obj, _fl(out)))
        return out

    def warm_start(self, net, gate=True, smoothing=0):
        """Start the open substep from the network's prediction (after `smoothing` one-ring averaging
        passes); gate: False / True (gate now) / "deferred" (the next iterate_to_tolerance's first check
        is the gate); returns (accepted, rho) per object (rho NaN when deferred)."""
        mode = 2 if gate == "deferred" else int(bool(gate))
        acc = np.zeros(self.n_objects, dtype=np.int32)
        rho = np.zeros(self.n_objects)
        self._check(self._L.spg_vbd_warm_start(self.h, net.h, int(smoothing), mode, _i(acc), _d(rho)))
        return acc.astype(bool), rho

    def history_start(self, gate=True):
        """Start the open substep from x~ + (x* - x~) of the previous substep, with the gates of warm_start;
        returns (accepted, rho) per object (rho NaN when deferred)."""
        mode = 2 if gate == "deferred" else int(bool(gate))
        acc = np.zeros(self.n_objects, dtype=np.int32)
        rho = np.zeros(self.n_objects)
        self._check(self._L.spg_vbd_history_start(self.h, mode, _i(acc), _d(rho)))
        return acc.astype(bool), rho

    def start_rho(self):
        """Per object |r(start)| / |r(x~)| of the substep's start as its gate measured it (NaN: not measured)."""
        rho = np.zeros(self.n_objects)
        self._check(self._L.spg_vbd_start_rho(self.h, _d(rho)))
        return rho

    def launch(self):
        c = VbdLaunch()
        self._check(self._L.spg_vbd_get_launch(self.h, ctypes.byref(c)))
        return c.as_dict()

    def set_launch(self, **kw):
        c = VbdLaunch()
        self._check(self._L.spg_vbd_get_launch(self.h, ctypes.byref(c)))
        for k, v in kw.items():
            setattr(c, k, int(v))
        self._check(self._L.spg_vbd_set_launch(self.h, ctypes.byref(c)))

    def autotune(self, trials=3):
        c = VbdLaunch()
        self._check(self._L.spg_vbd_autotune(self.h, trials, ctypes.byref(c)))
        return c.as_dict()

    def stats(self):
        s = VbdStats()
        self._check(self._L.spg_vbd_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._check(self._L.spg_vbd_reset_stats(self.h))

    def set_timing(self, enabled):
        self._check(self._L.spg_vbd_set_timing(self.h, int(enabled)))

    def synchronize(self):
        self._check(self._L.spg_vbd_synchronize(self.h))


def color_vertices(n_vertices, sets):
    """coloring::colorVertices over [(flat stencils, stencil size)]."""
    L = vbd_lib()
    arrs = [_i32(s).ravel() for s, _ in sets]
    ptrs = (_ip * max(1, len(arrs)))(*[_i(a) for a in arrs])
    lens = np.array([a.size for a in arrs] or [0], dtype=np.int64)
    sizes = np.array([k for _, k in sets] or [1], dtype=np.int32)
    out = np.zeros(n_vertices, dtype=np.int32)
    rc = L.spg_vbd_color_vertices(n_vertices, len(arrs), ptrs, lens.ctypes.data_as(_lp), _i(sizes), _i(out))
    if rc != 0:
        raise NativeError(rc, L.spg_vbd_last_error(None).decode())
    return out


def bench_copy(nbytes, repeats=10, device=0):
    """Mean time (ms) of a plain float4 copy of nbytes (calibration kernel for traffic counters)."""
    L = vbd_lib()
    ms = ctypes.c_double(0.0)
    rc = L.spg_vbd_bench_copy(int(nbytes), int(repeats), device, ctypes.byref(ms))
    if rc != 0:
        raise NativeError(rc, L.spg_vbd_last_error(None).decode())
    return ms.value


def bench_barrier(grid, block=256, group=None, rounds=500, device=0):
    """(us per barrier round inside one cooperative launch of grid x block threads, counted in groups of `group`
    blocks (default: the whole grid), us per back-to-back empty launch of the same shape); None for the barrier
    when the runtime refuses the cooperative launch."""
    L = vbd_lib()
    b, l = ctypes.c_double(0.0), ctypes.c_double(0.0)
    rc = L.spg_vbd_bench_barrier(int(grid), int(block), int(group or grid), int(rounds), device, ctypes.byref(b),
                                 ctypes.byref(l))
    if rc == 2:
        return None, None
    if rc != 0:
        raise NativeError(rc, "spg_vbd_bench_barrier: invalid arguments")
    return (b.value if b.value >= 0 else None), l.value


def element_derivatives(kind, xs, params, precision=0, device=0, form=0):
    """Per-vertex gradients (n, N, 3) and diagonal Hessian blocks (n, N, 3, 3) from the solver's closed
    forms, on GPU `device` or, with device=-1, their host instantiation (derivative checks)."""
    L = vbd_lib()
    t = _kind(kind)
    ns = STENCIL_SIZE[t]
    xs = _f64(xs, (-1, ns * 3))
    n = xs.shape[0]
    p = _f64(params, (n, PARAMS[t]))
    g = np.zeros((n, ns, 3))
    h = np.zeros((n, ns, 3, 3))
    rc = L.spg_vbd_element_derivatives(t, n, _d(xs), _d(p), precision, form, device, _d(g), _d(h))
    if rc != 0:
        raise NativeError(rc, L.spg_vbd_last_error(None).decode())
    return g, h


# ------------------------------------------------------------------------------------------------
# Learned warm start network (include/spg_vbd_learned.h)
# ------------------------------------------------------------------------------------------------

MLP_FEATURES, MLP_HIDDEN, MLP_OUTPUTS, MLP_NUM_PARAMS = 24, 128, 3, 36611
MLP_SHAPES = [(128, 24), (128, 128), (128, 128), (3, 128)]  # (out, in) of each layer


class Mlp:
    """spg_mlp: the per-vertex network, trained and evaluated by the library's own GPU kernels."""

    def __init__(self, device=0, seed=1):
        self._L = vbd_lib()
        h = ctypes.c_void_p()
        rc = self._L.spg_mlp_create(ctypes.byref(h), device, seed)
        if rc != 0:
            raise NativeError(rc, self._L.spg_mlp_last_error(None).decode())
        self.h = h.value
        self.device = device

    def __del__(self):
        if getattr(self, "h", None):
            self._L.spg_mlp_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise NativeError(rc, self._L.spg_mlp_last_error(self.h).decode())

    def params(self):
        out = np.zeros(MLP_NUM_PARAMS, dtype=np.float32)
        self._check(self._L.spg_mlp_get_params(self.h, _fl(out)))
        return out

    def set_params(self, p):
        p = _f32(p)
        if p.size != MLP_NUM_PARAMS:
            raise ValueError(f"expected {MLP_NUM_PARAMS} parameters")
        self._check(self._L.spg_mlp_set_params(self.h, _fl(p)))

    def normalization(self):
        m = np.zeros(MLP_FEATURES, dtype=np.float32)
        s = np.zeros(MLP_FEATURES, dtype=np.float32)
        self._check(self._L.spg_mlp_get_normalization(self.h, _fl(m), _fl(s)))
        return m, s

    def set_normalization(self, mean, std):
        m, s = _f32(mean, (MLP_FEATURES,)), _f32(std, (MLP_FEATURES,))
        self._check(self._L.spg_mlp_set_normalization(self.h, _fl(m), _fl(s)))

    def set_history_base(self, enabled):
        """True: the output is the change of the previous substep's offset; the warm start adds it to the
        history start x~ + (x* - x~)_previous."""
        self._check(self._L.spg_mlp_set_history_base(self.h, int(bool(enabled))))

    def history_base(self):
        v = np.zeros(1, dtype=np.int32)
        self._check(self._L.spg_mlp_get_history_base(self.h, _i(v)))
        return bool(v[0])

    def forward(self, features):
        f = _f32(features, (-1, MLP_FEATURES))
        out = np.zeros((f.shape[0], MLP_OUTPUTS), dtype=np.float32)
        self._check(self._L.spg_mlp_forward(self.h, _fl(f), f.shape[0], _fl(out)))
        return out

    def loss_grad(self, features, targets):
        f = _f32(features, (-1, MLP_FEATURES))
        t = _f32(targets, (f.shape[0], MLP_OUTPUTS))
        loss = ctypes.c_double(0.0)
        g = np.zeros(MLP_NUM_PARAMS, dtype=np.float32)
        self._check(self._L.spg_mlp_loss_grad(self.h, _fl(f), _fl(t), f.shape[0], ctypes.byref(loss), _fl(g)))
        return loss.value, g

    def train(self, features, targets, epochs=10, batch=4096, lr=1e-3, seed=1, standardize=True):
        f = _f32(features, (-1, MLP_FEATURES))
        t = _f32(targets, (f.shape[0], MLP_OUTPUTS))
        losses = np.zeros(max(1, epochs))
        self._check(self._L.spg_mlp_train(self.h, _fl(f), _fl(t), f.shape[0], epochs, batch, lr, seed,
                                          int(bool(standardize)), _d(losses)))
        return losses[:epochs]

    def save(self, path):
        m, s = self.normalization()
        np.savez(path, params=self.params(), mean=m, std=s)

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])
            self.set_normalization(z["mean"], z["std"])


def mlp_unflatten(p):
    """Flat parameter vector -> [(W (out, in), b (out,)), ...] in layer order."""
    p = np.asarray(p)
    out, k = [], 0
    for o, i in MLP_SHAPES:
        w = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((w, b))
    assert k == p.size
    return out


# ------------------------------------------------------------------------------------------------
# XPBD C-ABI context (include/spg_xpbd.h)
# ------------------------------------------------------------------------------------------------

def inverse_masses(m):
    """ParticleGroup::invMasses() of masses m: 0 for zero mass, 1/m otherwise (particleGroup.cpp:81)."""
    m = _f64(m)
    w = np.zeros_like(m)
    nz = m != 0
    w[nz] = 1.0 / m[nz]
    return w


class XpbdContext:
    """A spg_xpbd_ctx: one or more particle groups advanced together by XPBD on one GPU."""

    def __init__(self, objects, parallel_gauss_seidel=True, device=0):
        """objects: dicts with x (n,3), v (n,3) or None, w (n,) inverse masses (or m (n,) masses),
        energies [(type, stencils, params in the spg_xpbd_energy layout)]."""
        self._L = vbd_lib()
        keep = []
        objs = (XpbdObject * len(objects))()
        for k, o in enumerate(objects):
            x = _f64(o["x"], (-1, 3))
            v = None if o.get("v") is None else _f64(o["v"], (-1, 3))
            w = _f64(o["w"]) if o.get("w") is not None else inverse_masses(o["m"])
            ens = (XpbdEnergy * max(1, len(o["energies"])))()
            for j, (t, st, prm) in enumerate(o["energies"]):
                t = _kind(t)
                st = _i32(st, (-1, STENCIL_SIZE[t]))
                prm = _f64(prm, (st.shape[0], XPBD_PARAMS[t]))
                ens[j] = XpbdEnergy(t, st.shape[0], _i(st), _d(prm))
                keep += [st, prm]
            objs[k] = XpbdObject(x.shape[0], _d(x), _d(v), _d(w), len(o["energies"]), ens)
            keep += [x, v, w, ens]
        h = ctypes.c_void_p()
        rc = self._L.spg_xpbd_create(ctypes.byref(h), objs, len(objects), int(parallel_gauss_seidel), device)
        if rc != 0:
            raise NativeError(rc, self._L.spg_xpbd_last_error(None).decode())
        self.h = h.value
        self._owned = True
        self._owner = None
        self.n_objects = len(objects)
        self._n = [int(_f64(o["x"]).size // 3) for o in objects]
        self._ne = [len(o["energies"]) for o in objects]

    @classmethod
    def _borrow(cls, ptr, owner):
        self = cls.__new__(cls)
        self._L = vbd_lib()
        self.h = ptr
        self._owned = False
        self._owner = owner
        n = ctypes.c_int32(0)
        self._L.spg_xpbd_num_objects(ptr, ctypes.byref(n))
        self.n_objects = n.value
        self._n = []
        for o in range(self.n_objects):
            k = ctypes.c_int64(0)
            self._L.spg_xpbd_num_particles(ptr, o, ctypes.byref(k))
            self._n.append(k.value)
        return self

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "h", None):
            self._L.spg_xpbd_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise NativeError(rc, self._L.spg_xpbd_last_error(self.h).decode())

    def num_particles(self, obj=0):
        return self._n[obj]

    def num_groups(self, obj, energy):
        n = ctypes.c_int32(0)
        self._check(self._L.spg_xpbd_num_groups(self.h, obj, energy, ctypes.byref(n)))
        return n.value

    def groups(self, obj, energy, n_stencils):
        out = np.zeros(n_stencils, dtype=np.int32)
        self._check(self._L.spg_xpbd_get_groups(self.h, obj, energy, _i(out)))
        return out

    def num_launches(self):
        n = ctypes.c_int32(0)
        self._check(self._L.spg_xpbd_num_launches(self.h, ctypes.byref(n)))
        return n.value

    def state(self, obj=0):
        x, v = np.zeros((self._n[obj], 3)), np.zeros((self._n[obj], 3))
        self._check(self._L.spg_xpbd_get_state(self.h, obj, _d(x), _d(v)))
        return x, v

    def set_state(self, obj, x=None, v=None):
        self._check(self._L.spg_xpbd_set_state(self.h, obj, _d(None if x is None else _f64(x, (-1, 3))),
                                               _d(None if v is None else _f64(v, (-1, 3)))))

    def step(self, dt, nsubsteps=1, gravity=(0.0, -9.8, 0.0)):
        g = _f64(gravity)
        self._check(self._L.spg_xpbd_step(self.h, dt, nsubsteps, _d(g)))

    def save_state(self):
        self._check(self._L.spg_xpbd_save_state(self.h))

    def restore_state(self):
        self._check(self._L.spg_xpbd_restore_state(self.h))

    def launch(self):
        cfg = XpbdLaunch()
        self._check(self._L.spg_xpbd_get_launch(self.h, ctypes.byref(cfg)))
        return cfg.as_dict()

    def set_launch(self, **kw):
        cur = self.launch()
        cur.update(kw)
        cfg = XpbdLaunch(**cur)
        self._check(self._L.spg_xpbd_set_launch(self.h, ctypes.byref(cfg)))

    def stats(self):
        st = XpbdStats()
        self._check(self._L.spg_xpbd_get_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        self._check(self._L.spg_xpbd_reset_stats(self.h))

    def set_timing(self, enabled):
        self._check(self._L.spg_xpbd_set_timing(self.h, int(enabled)))

    def synchronize(self):
        self._check(self._L.spg_xpbd_synchronize(self.h))


def xpbd_color_stencils(stencils):
    """coloring::colorStencils through the C-ABI (host computation)."""
    L = vbd_lib()
    st = _i32(stencils)
    if st.ndim != 2:
        raise ValueError("stencils must be (n, size)")
    out = np.zeros(st.shape[0], dtype=np.int32)
    rc = L.spg_xpbd_color_stencils(st.shape[0], st.shape[1], _i(st), _i(out))
    if rc != 0:
        raise NativeError(rc, L.spg_xpbd_last_error(None).decode())
    return out


def _load_mesh(fn, path, per):
    L = host_lib()
    nv, nc = ctypes.c_int64(0), ctypes.c_int64(0)
    p = os.fsencode(path)
    _host_check(fn(p, ctypes.byref(nv), ctypes.byref(nc), None, None))
    x = np.zeros((nv.value, 3))
    c = np.zeros((nc.value, per), dtype=np.int32)
    _host_check(fn(p, ctypes.byref(nv), ctypes.byref(nc), _d(x), _i(c)))
    return x, c


def load_obj(path):
    """spg::io::loadObj: (vertices (n, 3), triangles (m, 3)) of an OBJ file's first shape."""
    return _load_mesh(host_lib().spg_io_load_obj, path, 3)


def load_msh(path):
    """spg::io::loadMsh: (nodes (n, 3), tetrahedra (m, 4)) of an ASCII MSH 4.1 file."""
    return _load_mesh(host_lib().spg_io_load_msh, path, 4)


def implicit_element_derivatives(kind, xs, params, device=0):
    """Gradients (n, 3 size) and Hessians (n, 3 size, 3 size) of n elements (xs (n, 3 size), params in the
    spg_vbd_energy layout) by the implicit solvers' closed forms: device -1 on the host, else the GPU kernel."""
    L = vbd_lib()
    t = _kind(kind)
    ns = STENCIL_SIZE[t]
    xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3 * ns)
    p = np.ascontiguousarray(params, dtype=np.float64).reshape(xs.shape[0], PARAMS[t])
    n = xs.shape[0]
    g = np.zeros((n, 3 * ns))
    h = np.zeros((n, 3 * ns, 3 * ns))
    rc = L.spg_implicit_element_derivatives(t, n, _d(xs), _d(p), _d(g), _d(h), int(device))
    if rc:
        raise NativeError(rc, L.spg_implicit_last_error(None).decode())
    return g, h


def xpbd_project_stencils(kind, xs, inv_masses, params, dt, precision=0, device=0):
    """n independent projections on the GPU by the projection kernels' device functions: (n, size, 3) dx."""
    L = vbd_lib()
    t = _kind(kind)
    ns = STENCIL_SIZE[t]
    xs = _f64(xs, (-1, ns * 3))
    n = xs.shape[0]
    w = _f64(inv_masses, (n, ns))
    p = _f64(params, (n, XPBD_PARAMS[t]))
    out = np.zeros((n, ns, 3))
    rc = L.spg_xpbd_project_stencils(t, n, _d(xs), _d(w), _d(p), dt, precision, device, _d(out))
    if rc != 0:
        raise NativeError(rc, L.spg_xpbd_last_error(None).decode())
    return out


# ------------------------------------------------------------------------------------------------
# Implicit Euler family (include/spg_implicit.h)
# ------------------------------------------------------------------------------------------------

BARAFF_WITKIN, NEWTON_DV, NEWTON_DX, NEWTON_ROBUST, BDF2 = 0, 1, 2, 3, 4
IMPLICIT_SOLVERS = {"baraff_witkin": BARAFF_WITKIN, "newton_dv": NEWTON_DV, "newton_dx": NEWTON_DX,
                    "newton_robust": NEWTON_ROBUST, "bdf2": BDF2}


class ImplicitContext:
    """A spg_implicit_ctx: particle groups advanced together by one implicit solver (one global system)."""

    def __init__(self, objects, solver, device=0):
        """objects: the dicts of VbdContext (x, v, m, energies in the stiffness-form spg_vbd_energy layout)."""
        self._L = vbd_lib()
        solver = IMPLICIT_SOLVERS[solver] if isinstance(solver, str) else int(solver)
        keep = []
        objs = (VbdObject * len(objects))()
        for k, o in enumerate(objects):
            x = _f64(o["x"], (-1, 3))
            v = None if o.get("v") is None else _f64(o["v"], (-1, 3))
            m = _f64(o["m"])
            ens = (VbdEnergy * max(1, len(o["energies"])))()
            for j, (t, st, prm) in enumerate(o["energies"]):
                t = _kind(t)
                st = _i32(st, (-1, STENCIL_SIZE[t]))
                prm = _f64(prm, (st.shape[0], PARAMS[t]))
                ens[j] = VbdEnergy(t, st.shape[0], _i(st), _d(prm))
                keep += [st, prm]
            objs[k] = VbdObject(x.shape[0], _d(x), _d(v), _d(m), len(o["energies"]), ens, None)
            keep += [x, v, m, ens]
        h = ctypes.c_void_p()
        rc = self._L.spg_implicit_create(ctypes.byref(h), objs, len(objects), solver, device)
        if rc != 0:
            raise NativeError(rc, self._L.spg_implicit_last_error(None).decode())
        self.h = h.value
        self.solver = solver
        self.n_objects = len(objects)
        self._n = [int(_f64(o["x"]).size // 3) for o in objects]

    def __del__(self):
        if getattr(self, "h", None):
            self._L.spg_implicit_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise NativeError(rc, self._L.spg_implicit_last_error(self.h).decode())

    def num_dofs(self):
        n = ctypes.c_int64(0)
        self._check(self._L.spg_implicit_num_dofs(self.h, ctypes.byref(n)))
        return n.value

    def num_blocks(self):
        n = ctypes.c_int64(0)
        self._check(self._L.spg_implicit_num_blocks(self.h, ctypes.byref(n)))
        return n.value

    def state(self, obj=0):
        x, v = np.zeros((self._n[obj], 3)), np.zeros((self._n[obj], 3))
        self._check(self._L.spg_implicit_get_state(self.h, obj, _d(x), _d(v)))
        return x, v

    def set_state(self, obj, x=None, v=None):
        self._check(self._L.spg_implicit_set_state(self.h, obj, _d(None if x is None else _f64(x, (-1, 3))),
                                                   _d(None if v is None else _f64(v, (-1, 3)))))

    def step(self, dt, nsubsteps=1, gravity=(0.0, -9.8, 0.0)):
        g = _f64(gravity)
        self._check(self._L.spg_implicit_step(self.h, dt, nsubsteps, _d(g)))

    def reset(self):
        self._check(self._L.spg_implicit_reset(self.h))

    def system_force(self, gravity):
        g = _f64(gravity)
        f = np.zeros(self.num_dofs())
        self._check(self._L.spg_implicit_system_force(self.h, _d(g), _d(f)))
        return f

    def stiffness(self):
        """K = -Hessian at the current positions as scipy.sparse.bsr_matrix (3x3 blocks)."""
        import scipy.sparse as sp
        nb = self.num_blocks()
        nrows = self.num_dofs() // 3
        ptr = np.zeros(nrows + 1, dtype=np.int64)
        cols = np.zeros(nb, dtype=np.int32)
        blocks = np.zeros((nb, 3, 3))
        self._check(self._L.spg_implicit_stiffness(self.h, ptr.ctypes.data_as(_lp), _i(cols), _d(blocks)))
        return sp.bsr_matrix((blocks, cols, ptr), shape=(3 * nrows, 3 * nrows))

    def cg_solve(self, alpha, beta, b, tol=1e-5, max_iters=300):
        b = _f64(b)
        x = np.zeros(self.num_dofs())
        it = ctypes.c_int32(0)
        err = ctypes.c_double(0)
        self._check(self._L.spg_implicit_cg_solve(self.h, alpha, beta, _d(b), tol, max_iters, _d(x),
                                                  ctypes.byref(it), ctypes.byref(err)))
        return x, it.value, err.value

    def bench_spmv(self, alpha, beta, repeats=20):
        """Mean ms of one block-sparse product with alpha M + beta K (measurement hook)."""
        ms = ctypes.c_double(0)
        self._check(self._L.spg_implicit_bench_spmv(self.h, alpha, beta, repeats, ctypes.byref(ms)))
        return ms.value

    def spmv_bytes(self):
        """Compulsory bytes of one product of the upper-block storage: 72 B values + 4 B column per upper
        block, 4 B index + 4 B column per lower block (read transposed from the upper one), 24 B of x and
        24 B of y and 16 B of row pointers per block row."""
        n = self.num_dofs() // 3
        n_upper = (self.num_blocks() + n) // 2
        return 76.0 * n_upper + 8.0 * (n_upper - n) + 48.0 * n + 16.0 * (n + 1)

    def stats(self):
        st = ImplicitStats()
        self._check(self._L.spg_implicit_get_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        self._check(self._L.spg_implicit_reset_stats(self.h))

    def set_timing(self, enabled):
        self._check(self._L.spg_implicit_set_timing(self.h, int(enabled)))

    def synchronize(self):
        self._check(self._L.spg_implicit_synchronize(self.h))


def _rigid_kind(k):
    return RIGID_KIND.get(k, k)


def rigid_element_derivatives(kind, com, rot, params, device=0):
    """Closed-form gradient (n, 6 s) and per-body diagonal Hessian blocks (n, s, 6, 6) of rigid-body elements at
    theta = 0; device -1 evaluates the same code on the host."""
    t = _rigid_kind(kind)
    ns, np_ = RIGID_STENCIL[t], RIGID_PARAMS[t]
    com = _f64(com, (-1, 3 * ns))
    n = com.shape[0]
    rot = _f64(rot, (n, 9 * ns))
    prm = _f64(params, (n, np_))
    g = np.zeros((n, 6 * ns))
    h = np.zeros((n, ns, 6, 6))
    L = vbd_lib()
    rc = L.spg_rigid_element_derivatives(t, n, _d(com), _d(rot), _d(prm), _d(g), _d(h), device)
    if rc != 0:
        raise NativeError(rc, L.spg_rigid_last_error(None).decode())
    return g, h


class RigidContext:
    """A spg_rigid_ctx: RigidBodyGroups advanced together by VBD on one GPU (FP64)."""

    def __init__(self, groups, device=0, graph=True):
        """groups: objects with x, theta (nb, 3), m (nb,), inertia (nb, 3, 3), energies [(kind, stencils, args)]
        and optionally v, omega (a scenes.RigidScene), or dicts with the same keys."""
        self._L = vbd_lib()
        get = (lambda o, k: o.get(k)) if isinstance(groups[0], dict) else (lambda o, k: getattr(o, k, None))
        keep = []
        gs = (RigidGroup * len(groups))()
        self._nb = []
        for k, o in enumerate(groups):
            x = _f64(get(o, "x"), (-1, 3))
            nb = x.shape[0]
            th = _f64(get(o, "theta"), (nb, 3))
            m = _f64(get(o, "m"), (nb,))
            inertia = _f64(get(o, "inertia"), (nb, 9))
            v = get(o, "v")
            w = get(o, "omega")
            v = None if v is None else _f64(v, (nb, 3))
            w = None if w is None else _f64(w, (nb, 3))
            energies = get(o, "energies")
            ens = (RigidEnergy * max(1, len(energies)))()
            for j, (kind, st, prm) in enumerate(energies):
                t = _rigid_kind(kind)
                st = _i32(st, (-1, RIGID_STENCIL[t]))
                prm = _f64(prm, (st.shape[0], RIGID_PARAMS[t]))
                ens[j] = RigidEnergy(t, st.shape[0], _i(st), _d(prm))
                keep += [st, prm]
            gs[k] = RigidGroup(nb, _d(x), _d(th), _d(v) if v is not None else None,
                               _d(w) if w is not None else None, _d(m), _d(inertia), len(energies), ens)
            keep += [x, th, m, inertia, v, w, ens]
            self._nb.append(nb)
        h = ctypes.c_void_p()
        rc = self._L.spg_rigid_create(ctypes.byref(h), gs, len(groups), device)
        if rc != 0:
            raise NativeError(rc, self._L.spg_rigid_last_error(None).decode())
        self.h = h.value
        if not graph:
            self._L.spg_rigid_set_graph(self.h, 0)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.spg_rigid_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise NativeError(rc, self._L.spg_rigid_last_error(self.h).decode())

    def step(self, dt, nsubsteps=1, iterations=10, gravity=(0.0, -9.81, 0.0)):
        g = np.asarray(gravity, dtype=np.float64)
        self._check(self._L.spg_rigid_step(self.h, dt, nsubsteps, iterations, _d(g)))

    def state(self, group=0):
        """(x, theta, v, omega), each (nb, 3)."""
        out = [np.zeros((self._nb[group], 3)) for _ in range(4)]
        self._check(self._L.spg_rigid_get_state(self.h, group, *[_d(a) for a in out]))
        return tuple(out)

    def set_state(self, group=0, x=None, theta=None, v=None, omega=None):
        nb = self._nb[group]
        arrs = [None if a is None else _f64(a, (nb, 3)) for a in (x, theta, v, omega)]
        self._check(self._L.spg_rigid_set_state(self.h, group, *[None if a is None else _d(a) for a in arrs]))

    def rotations(self, group=0):
        """rotationMatrices() of the group, (nb, 9) row-major."""
        R = np.zeros((self._nb[group], 9))
        self._check(self._L.spg_rigid_get_rotations(self.h, group, _d(R)))
        return R

    def set_rotations(self, R, group=0):
        R = _f64(R, (self._nb[group], 9))
        self._check(self._L.spg_rigid_set_rotations(self.h, group, _d(R)))

    def colors(self, group=0):
        c = np.zeros(self._nb[group], dtype=np.int32)
        n = ctypes.c_int32(0)
        self._check(self._L.spg_rigid_get_colors(self.h, group, _i(c), ctypes.byref(n)))
        return c, n.value

    def stats(self):
        st = RigidStats()
        self._check(self._L.spg_rigid_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in RigidStats._fields_}

    def profile_substep(self, h, iterations, gravity=(0.0, -9.81, 0.0)):
        """One eager substep with HIP events around each colour launch: (summed ms, launches)."""
        g = np.asarray(gravity, dtype=np.float64)
        ms = ctypes.c_double(0)
        n = ctypes.c_int32(0)
        self._check(self._L.spg_rigid_profile_substep(self.h, h, iterations, _d(g), ctypes.byref(ms),
                                                      ctypes.byref(n)))
        return ms.value, n.value
