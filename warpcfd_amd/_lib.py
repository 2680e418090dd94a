"""ctypes binding of libwcfd.so (C ABI declared in include/wcfd.h).

The product path has no CPU fallback: if the library is missing or no GPU is
present, the first call that needs it raises.
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = 'libwcfd.so'
LIB_PATH = os.path.join(_HERE, LIB_NAME)
# the same library built with double fields (-DWCFD_REAL=double)
LIB64_NAME = 'libwcfd64.so'
LIB64_PATH = os.path.join(_HERE, LIB64_NAME)
HEADER_PATH = os.path.join(os.path.dirname(_HERE), 'include', 'wcfd.h')

WCFD_OK, WCFD_EINVAL, WCFD_EHIP, WCFD_ENOTCONV, WCFD_ENAN, WCFD_ESCRATCH = 0, 1, 2, 3, 4, 5
WCFD_EMESH, WCFD_ECAPACITY = 6, 7
NOALLOC_MAX_BLOCKS = 16384
DIRICHLET, NEUMANN = 1, 2
SCHEME_CORRECTED, SCHEME_IMPLICIT_OUTFLOW = 1, 2
CG_STANDARD, CG_SINGLE_REDUCTION = 0, 1
CONV_UPWIND, CONV_CENTRAL, CONV_UPWIND_LINEAR = 0, 1, 2
MONITOR_LEN = 48

c_p = ctypes.c_void_p
c_i32 = ctypes.c_int32
c_f = ctypes.c_float
c_d = ctypes.c_double


class WcfdMesh(ctypes.Structure):
    _fields_ = [('n_cells', c_i32), ('n_faces', c_i32), ('n_internal', c_i32),
                ('n_boundary', c_i32), ('nnz', c_i32), ('n_slots', c_i32),
                ('owner', c_p), ('neigh', c_p), ('psi', c_p), ('skew', c_p), ('area', c_p),
                ('normal', c_p), ('dist', c_p), ('dmag', c_p), ('volume', c_p),
                ('slot_face', c_p), ('slot_side', c_p), ('slot_sn', c_p), ('slot_pos', c_p),
                ('diag_pos', c_p), ('row_ptr', c_p), ('col', c_p), ('slot_nbr', c_p),
                ('slot_dnb', c_p)]


class WcfdBC(ctypes.Structure):
    _fields_ = [('type', c_p), ('value', c_p)]


class SolveInfo(ctypes.Structure):
    _fields_ = [('iterations', c_i32), ('converged', c_i32), ('residual', c_d),
                ('threshold', c_d)]

    def as_tuple(self):
        return (self.iterations, self.residual, self.threshold)


class WcfdBatch(ctypes.Structure):
    _fields_ = [('n_cases', c_i32), ('n_blocks', c_i32), ('chunk', c_i32), ('block', c_i32),
                ('cell_off', c_p), ('cell_end', c_p), ('blk_off', c_p), ('blk_case', c_p), ('cell_case', c_p),
                ('face_case', c_p), ('active', c_p), ('nu', c_p), ('ref_cell', c_p),
                ('ref_value', c_p), ('stop', c_p), ('monitor', c_p), ('outer_iters', c_p),
                ('max_case_blocks', c_i32), ('n_cell_blk', c_i32), ('n_face_blk', c_i32),
                ('n_bface_blk', c_i32), ('cell_blk', c_p), ('face_blk', c_p), ('bface_blk', c_p)]


class SimpleParams(ctypes.Structure):
    _fields_ = [('nu', c_f), ('density', c_f), ('alpha_u', c_f), ('alpha_p', c_f), ('dt', c_f),
                ('scheme_flags', c_i32), ('ref_cell', c_i32), ('ref_value', c_f),
                ('n_inner', c_i32), ('n_ortho', c_i32), ('atol', c_d), ('tol', c_d), ('rtol0', c_d),
                ('maxiter', c_i32), ('check_every', c_i32), ('first_iteration', c_i32),
                ('cg_variant', c_i32), ('check_every_u', c_i32), ('p_refine', c_i32),
                ('graph_groups_u', c_i32), ('graph_groups_p', c_i32), ('maxiter_u', c_i32)]


class SimpleState(ctypes.Structure):
    _fields_ = [(n, c_p) for n in ('cell_values', 'cell_grad', 'face_values', 'mass_flux', 'hbya',
                                   'p_x', 'rua', 'rua_face', 'mom_values', 'mom_rhs', 'p_values',
                                   'p_rhs', 'resid')]


class SimpleInfo(ctypes.Structure):
    _fields_ = [('momentum', SolveInfo), ('pressure', SolveInfo), ('continuity', c_d),
                ('velocity_change', c_d)]


_PM = ctypes.POINTER(WcfdMesh)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PB = ctypes.POINTER(WcfdBatch)
_SIGS = {
    'wcfd_last_hip_error': (ctypes.c_int, []),
    'wcfd_error_string': (ctypes.c_char_p, [ctypes.c_int]),
    'wcfd_abi_version': (ctypes.c_int, []),
    'wcfd_real_bytes': (ctypes.c_int, []),
    'wcfd_set_boundary_conditions': (ctypes.c_int, [_PM, WcfdBC, c_p, ctypes.c_int, c_p]),
    'wcfd_face_interpolation': (ctypes.c_int, [_PM, WcfdBC, c_p, c_p, c_p, ctypes.c_int,
                                               ctypes.c_int, c_p, c_p, c_p, c_p]),
    'wcfd_face_interpolation_upwind': (ctypes.c_int, [_PM, WcfdBC, c_p, c_p, c_p, c_p,
                                                      ctypes.c_int, ctypes.c_int, c_p]),
    'wcfd_green_gauss': (ctypes.c_int, [_PM, c_p, c_p, ctypes.c_int, ctypes.c_int, c_p]),
    'wcfd_mass_flux': (ctypes.c_int, [_PM, c_p, c_p, c_p]),
    'wcfd_divergence': (ctypes.c_int, [_PM, c_p, c_p, c_p]),
    'wcfd_assemble_momentum': (ctypes.c_int, [_PM, WcfdBC, c_p, c_p, c_p, c_p, ctypes.c_int, c_f,
                                              c_f, c_f, c_f, ctypes.c_int, c_p, c_p, c_p]),
    'wcfd_assemble_pressure': (ctypes.c_int, [_PM, WcfdBC, c_p, c_p, c_p, c_p, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, c_f, c_p, c_p, c_p, c_p]),
    'wcfd_ell_to_csr': (ctypes.c_int, [_PM, c_p, c_p, c_p]),
    'wcfd_convection_weights': (ctypes.c_int, [_PM, WcfdBC, c_p, c_p, c_p, c_p, ctypes.c_int,
                                               ctypes.c_int, _PI32, ctypes.c_int, c_p, c_p]),
    'wcfd_diffusion_weights': (ctypes.c_int, [_PM, WcfdBC, c_p, c_p, c_p, c_f, ctypes.c_int,
                                              ctypes.c_int, c_f, _PI32, ctypes.c_int, c_p, c_p]),
    'wcfd_ddt_weights': (ctypes.c_int, [_PM, c_p, c_f, _PI32, ctypes.c_int, c_p, c_p]),
    'wcfd_grad_weights': (ctypes.c_int, [_PM, c_p, ctypes.c_int, c_p, c_p]),
    'wcfd_equation_add_face_term': (ctypes.c_int, [_PM, c_p, ctypes.c_int, c_f, c_p, c_p, c_p]),
    'wcfd_equation_add_cell_term': (ctypes.c_int, [_PM, c_p, ctypes.c_int, c_f, c_p, c_p, c_p]),
    'wcfd_equation_add_rhs': (ctypes.c_int, [ctypes.c_int32, c_p, c_f, c_p, c_p]),
    'wcfd_equation_relax': (ctypes.c_int, [_PM, c_p, _PI32, ctypes.c_int, c_f, c_p, c_p, c_p]),
    'wcfd_equation_replace_row': (ctypes.c_int, [_PM, ctypes.c_int, ctypes.c_int, ctypes.c_int32,
                                                 c_f, c_p, c_p, c_p]),
    'wcfd_spmv_ell': (ctypes.c_int, [_PM, c_p, ctypes.c_int, c_p, c_p, c_p]),
    'wcfd_cg_ell': (ctypes.c_int, [_PM, c_p, ctypes.c_int, c_p, c_p, c_d, c_d, ctypes.c_int,
                                   ctypes.c_int, c_p, ctypes.POINTER(SolveInfo), c_p]),
    'wcfd_bicgstab_ell': (ctypes.c_int, [_PM, c_p, ctypes.c_int, c_p, c_p, c_d, c_d, ctypes.c_int,
                                         ctypes.c_int, c_p, ctypes.POINTER(SolveInfo), c_p]),
    'wcfd_rua_hbya': (ctypes.c_int, [_PM, c_p, c_f, c_p, c_p, c_p, c_p, c_p]),
    'wcfd_relax_pressure': (ctypes.c_int, [_PM, c_p, c_f, c_p, c_p]),
    'wcfd_correct_velocity': (ctypes.c_int, [_PM, WcfdBC, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                             c_p, c_p]),
    'wcfd_reduce_workspace_bytes': (ctypes.c_size_t, [c_i32]),
    'wcfd_convergence_report': (ctypes.c_int, [_PM, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                               c_p]),
    'wcfd_convergence_workspace_bytes': (ctypes.c_size_t, []),
    'wcfd_batch_convergence_report': (ctypes.c_int, [_PM, ctypes.POINTER(WcfdBatch), c_p, c_p, c_p,
                                                     c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    'wcfd_batch_convergence_workspace_bytes': (ctypes.c_size_t, [c_i32]),
    'wcfd_batch_convergence_report_noalloc': (ctypes.c_int, [_PM, ctypes.POINTER(WcfdBatch), c_p,
                                                             c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                                             c_p]),
    'wcfd_spmv': (ctypes.c_int, [c_i32, c_p, c_p, c_p, ctypes.c_int, c_p, c_p, c_p]),
    'wcfd_dot_f64': (ctypes.c_int, [c_i32, c_p, c_p, c_p, c_p, c_p]),
    'wcfd_krylov_workspace_bytes': (ctypes.c_size_t, [c_i32, ctypes.c_int]),
    'wcfd_cg': (ctypes.c_int, [c_i32, c_p, c_p, c_p, c_p, ctypes.c_int, c_p, c_p, c_d, c_d,
                               ctypes.c_int, ctypes.c_int, c_p, ctypes.POINTER(SolveInfo), c_p]),
    'wcfd_bicgstab': (ctypes.c_int, [c_i32, c_p, c_p, c_p, c_p, ctypes.c_int, c_p, c_p, c_d, c_d,
                                     ctypes.c_int, ctypes.c_int, c_p, ctypes.POINTER(SolveInfo),
                                     c_p]),
    'wcfd_set_krylov_launch': (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    'wcfd_get_krylov_launch': (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    'wcfd_set_krylov_fold': (ctypes.c_int, [ctypes.c_int]),
    'wcfd_get_krylov_fold': (ctypes.c_int, []),
    'wcfd_set_krylov_inline_check': (ctypes.c_int, [ctypes.c_int]),
    'wcfd_get_krylov_inline_check': (ctypes.c_int, []),
    'wcfd_set_krylov_persistent': (ctypes.c_int, [ctypes.c_int]),
    'wcfd_get_krylov_persistent': (ctypes.c_int, []),
    'wcfd_krylov_persistent_stats': (ctypes.c_int, [ctypes.POINTER(ctypes.c_int64),
                                                     ctypes.POINTER(ctypes.c_int64)]),
    'wcfd_profile_timeline': (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    'wcfd_match_faces': (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int64]),
    'wcfd_debug_krylov_persistent_shrink': (ctypes.c_int, [ctypes.c_int]),
    'wcfd_simple_workspace_bytes': (ctypes.c_size_t, [c_i32, c_i32]),
    'wcfd_kernel_count': (ctypes.c_int, []),
    'wcfd_kernel_name': (ctypes.c_char_p, [ctypes.c_int]),
    'wcfd_profile_begin': (ctypes.c_int, [ctypes.c_int]),
    'wcfd_profile_end': (ctypes.c_int, [ctypes.POINTER(c_d), ctypes.POINTER(c_i32), ctypes.c_int]),
    'wcfd_profile_skipped': (ctypes.c_int, [ctypes.POINTER(c_i32), ctypes.c_int]),
    'wcfd_simple_iteration': (ctypes.c_int, [_PM, WcfdBC, ctypes.POINTER(SimpleParams),
                                             ctypes.POINTER(SimpleState), c_p,
                                             ctypes.POINTER(SimpleInfo), c_p]),
    'wcfd_monitor_update': (ctypes.c_int, [ctypes.POINTER(c_d), c_d, c_d, c_d, c_d, c_d]),
    'wcfd_batch_krylov_workspace_bytes': (ctypes.c_size_t, [c_i32, c_i32, c_i32]),
    'wcfd_batch_cg_ell': (ctypes.c_int, [_PM, _PB, c_p, ctypes.c_int, c_p, c_p, c_d, c_d,
                                         ctypes.c_int, ctypes.c_int, c_p, ctypes.POINTER(SolveInfo),
                                         c_p]),
    'wcfd_batch_bicgstab_ell': (ctypes.c_int, [_PM, _PB, c_p, ctypes.c_int, c_p, c_p, c_d, c_d,
                                               ctypes.c_int, ctypes.c_int, c_p,
                                               ctypes.POINTER(SolveInfo), c_p]),
    'wcfd_batch_simple_workspace_bytes': (ctypes.c_size_t, [c_i32, c_i32, c_i32, c_i32]),
    'wcfd_batch_simple_iteration': (ctypes.c_int, [_PM, _PB, WcfdBC, ctypes.POINTER(SimpleParams),
                                                   ctypes.POINTER(SimpleState), c_p,
                                                   ctypes.POINTER(SimpleInfo), c_p]),
}

_libs = {}


class WcfdError(RuntimeError):
    pass


def header_symbols(path=HEADER_PATH):
    """Function names declared in include/wcfd.h."""
    text = open(path).read()
    return sorted(set(re.findall(r'^[A-Za-z_][\w\s\*]*?\b(wcfd_\w+)\s*\(', text, re.M)))


def load(path=LIB_PATH):
    """Load libwcfd.so (or the library at ``path``) and declare every
    signature; raises if it is missing.  Each library is loaded once."""
    if path in _libs:
        return _libs[path]
    # torch first: it loads its own HIP runtime; libwcfd.so must bind to that
    # same runtime (loaded the other way round, HIP calls fail with
    # hipErrorNoDevice on tensors torch allocated)
    import torch  # noqa: F401
    if not os.path.exists(path):
        raise WcfdError(f'{os.path.basename(path)} is not built ({path}); run '
                        f'__graft_entry__.build() first')
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


def load64():
    """libwcfd64.so: the same entry points with double fields."""
    return load(LIB64_PATH)


def for_dtype(dtype):
    """The library whose field type is ``dtype`` (numpy or torch float32 /
    float64)."""
    import numpy as np
    import torch
    if dtype in (torch.float32, torch.float64):
        dtype = np.float32 if dtype == torch.float32 else np.float64
    dt = np.dtype(dtype)
    if dt == np.float32:
        return load()
    if dt == np.float64:
        return load64()
    raise ValueError(f'no library computes in {dt}: float32 or float64')


def real_dtype(lib):
    """torch dtype of a loaded library's fields."""
    import torch
    return torch.float64 if lib.wcfd_real_bytes() == 8 else torch.float32


def check(code, what=''):
    if code != WCFD_OK:
        lib = load()
        msg = lib.wcfd_error_string(code).decode()
        if code == WCFD_EHIP:
            msg += f' (hipError {lib.wcfd_last_hip_error()})'
        if code == WCFD_ENAN:
            raise ValueError(f'{what}: {msg}')
        raise WcfdError(f'{what}: {msg}')


class Profiler:
    """Per-kernel HIP-event timings of every libwcfd launch in a region."""

    def __init__(self, capacity=65536):
        self.capacity = capacity
        self.result = None
        self.skipped = {}

    def __enter__(self):
        check(load().wcfd_profile_begin(self.capacity), 'profile_begin')
        return self

    def __exit__(self, *exc):
        lib = load()
        n = lib.wcfd_kernel_count()
        ms = (c_d * n)()
        cnt = (c_i32 * n)()
        check(lib.wcfd_profile_end(ms, cnt, n), 'profile_end')
        skipped = (c_i32 * n)()
        check(lib.wcfd_profile_skipped(skipped, n), 'profile_skipped')
        self.result = {lib.wcfd_kernel_name(i).decode(): (ms[i], cnt[i]) for i in range(n) if cnt[i]}
        # launches that exited at entry (their solve had already stopped), left out above
        self.skipped = {lib.wcfd_kernel_name(i).decode(): skipped[i] for i in range(n) if skipped[i]}
        return False
