"""ctypes binding of ``libefea_hip.so`` (declarations: ``include/efea.h``).

The product path has no CPU fallback: if the library is missing, or no GPU
is visible, every entry point raises.
"""

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libefea_hip.so")

vp = C.c_void_p
i32, i64, f64, sz = C.c_int, C.c_int64, C.c_double, C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "efea_version": (i32, []),
    "efea_last_error": (C.c_char_p, []),
    "efea_elastic_ke": (i32, [i32, i32, i64, vp, vp, vp, i32, vp, i32, f64, i32, f64, vp, vp]),
    "efea_reaction_diffusion_ke": (
        i32,
        [i32, i32, i64, vp, vp, vp, i32, vp, i32, vp, f64, vp, i32, vp, f64, vp, vp, vp],
    ),
    "efea_boundary_load": (i32, [i32, i64, vp, vp, i32, vp, vp, vp]),
    "efea_elastic_ke_dense": (i32, [i32, i64, vp, vp, vp, i32, vp, i32, vp, i32, f64, vp, vp]),
    "efea_elastic_strain_stress": (i32, [i32, i64, vp, vp, vp, i32, vp, vp, i32, f64, i32, f64, vp, vp]),
    "efea_node_average": (i32, [i32, i32, vp, vp, vp, vp, vp]),
    "efea_csr_scatter": (i32, [i32, i64, vp, vp, vp, vp, vp]),
    "efea_cast_f64_f32": (i32, [i64, vp, vp, vp]),
    "efea_split_f32": (i32, [i32, i32, vp, vp, vp, vp, vp]),
    "efea_spmv_scratch_bytes": (sz, [i32, i32]),
    "efea_spmv": (i32, [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp]),
    "efea_spmv_repeat": (i32, [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp]),
    "efea_step1_bench_scratch_bytes": (sz, [i32, i32, i32]),
    "efea_step1_bench": (i32, [i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, i32, vp]),
    "efea_bj_setup": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp]),
    "efea_pcg_workspace_bytes": (sz, [i32, i32, i32]),
    "efea_default_group": (i32, [i32, i32, i32]),
    "efea_solve_pcg": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, f64, f64, f64, i32, i32, i32, i32, vp, vp],
    ),
    "efea_solve_pcg_noalloc": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, f64, f64, f64, i32, i32, i32, i32, vp, vp],
    ),
    "efea_coarse_ld": (i32, [i32, i32]),
    "efea_coarse_scratch_bytes": (sz, [i32]),
    "efea_coarse_setup": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, f64, vp, vp, vp]),
    "efea_pcg2_workspace_bytes": (sz, [i32, i32, i32]),
    "efea_solve_pcg2": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32, vp, vp, vp, sz, f64, f64, f64,
         i32, i32, i32, i32, vp, vp],
    ),
    "efea_pcg2p_workspace_bytes": (sz, [i32, i32, i32]),
    "efea_pcgp_workspace_bytes": (sz, [i32, i32]),
    "efea_pcgp_max_lanes": (i32, []),
    "efea_bicgstab_workspace_bytes": (sz, [i32, i32]),
    "efea_solve_bicgstab": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, sz, f64, i32, i32, vp, vp]),
    "efea_solve_pcgp": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, f64, f64, i32, i32, i32, i32, vp, vp],
    ),
    "efea_pcgp_start": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, f64, f64, i32, i32, i32, i32, vp],
    ),
    "efea_pcgp_finish": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, f64, f64, i32, i32, i32, i32, vp, vp, vp],
    ),
    "efea_pp_iter_bench": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, i32, vp, i32, vp, vp, i32, vp, vp, sz, i32, vp],
    ),
    "efea_solve_pcg2p": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32, i32, vp, vp, i32, vp, vp, vp,
         vp, sz, f64, f64, i32, i32, i32, vp, vp],
    ),
    "efea_pcg2p_start": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32, i32, vp, vp, i32, vp, vp, vp,
         vp, sz, f64, f64, i32, i32, i32, vp],
    ),
    "efea_pcg2p_finish": (
        i32,
        [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32, i32, vp, vp, i32, vp, vp, vp,
         vp, sz, f64, f64, i32, i32, i32, vp, vp, vp],
    ),
    "efea_batch_create": (i32, [i32, i32, i32, i32, i32, C.POINTER(vp)]),
    "efea_batch_destroy": (i32, [vp]),
    "efea_batch_add": (
        i32,
        [vp, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32, vp, vp, vp, sz, f64, f64, f64, i32, i32, vp],
    ),
    "efea_batch_run": (i32, [vp, i32, vp]),
    "efea_batch_launch": (i32, [vp, i32, vp]),
    "efea_batch_wait": (i32, [vp, vp]),
    "efea_batch_pop": (i32, [vp, C.POINTER(i32), vp]),
    "efea_batch_active": (i32, [vp]),
    "efea_dot": (i32, [i64, vp, vp, vp, vp, vp]),
    "efea_gather_sum": (i32, [i64, vp, vp, vp, vp]),
    "efea_pf_damage_ke": (
        i32,
        [i32, i64, vp, vp, vp, i32, vp, i32, vp, vp, vp, i32, f64, i32, vp, i32, f64, i32, i32, f64, vp, vp,
         vp, vp],
    ),
    "efea_pf_elastic_ke": (i32, [i32, i64, vp, vp, vp, i32, vp, vp, vp, i32, f64, i32, i32, f64, vp, vp]),
    "efea_pf_split_points": (i32, [i64, i32, vp, vp, i32, f64, i32, vp, vp, vp, vp, vp]),
    "efea_pf_strain_stress": (i32, [i32, i64, vp, vp, vp, i32, vp, vp, vp, i32, f64, i32, i32, f64, vp, vp]),
    "efea_uv_ke": (i32, [i32, i32, i64, vp, vp, vp, i32, vp, i32, f64, i32, f64, vp, vp]),
    "efea_lincomb": (i32, [i64, f64, vp, f64, vp, f64, vp, f64, vp, vp, vp]),
    "efea_time_update": (i32, [i32, i64, f64, f64, f64, f64, vp, vp, vp, vp, vp, vp, vp, vp]),
    "efea_zz1_energy": (i32, [i32, i64, vp, vp, vp, i32, vp, vp, f64, vp, vp]),
    "efea_bounded_lsq_workspace_bytes": (sz, [i32]),
    "efea_solve_bounded_lsq": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, sz, f64, f64, i32, i32, vp, vp]),
}

# time schemes (EFEA_SCHEME_* of include/efea.h)
SCHEME_CODE = {"elliptic": 0, "parabolic": 1, "newmark": 2, "midpoint": 3, "hht": 4, "hht_newmark": 5,
               "euler_implicit": 6, "euler_explicit": 7}

ELEM_CODE = {"TRI3": 0, "QUAD4": 1, "HEXA8": 2, "TETRA4": 3, "TETRA10": 4, "SEG2": 5}

ERR_NOCONV = -3

_lib = None


class EfeaError(RuntimeError):
    def __init__(self, name, code, msg):
        super().__init__(f"{name} failed ({code}): {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Loads the shared library (no GPU needed) and declares every symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback"
        )
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name: str, *args) -> int:
    """Calls ``name``; raises ``EfeaError`` on a negative status other than
    non-convergence, which is returned for the caller to handle."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if isinstance(rc, int) and rc < 0 and rc != ERR_NOCONV:
        raise EfeaError(name, rc, lib.efea_last_error().decode())
    return rc


def last_error() -> str:
    return load().efea_last_error().decode()
