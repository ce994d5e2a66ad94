"""ctypes binding of libphylo_hip.so (the C ABI in include/phylo_hip.h and
include/phylo_synth.h).

The library is built in-tree by ``phylocaml_amd/build.py`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared library is
missing, importing this module raises ImportError naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphylo_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} not found: build the HIP library first "
        "(python phylocaml_amd/build.py)"
    )

# torch brings its own HIP runtime; load it first so that the library binds to
# the same runtime instance as the tensors and streams it is handed (loaded
# the other way round, the library's HIP calls fail with no device).
try:
    import torch as _torch  # noqa: F401
except ImportError:  # host-only use (programs, synthetic data, oracle checks)
    _torch = None

lib = C.CDLL(LIB_PATH)

i32, i64, u64, f64, f32 = C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_float
vp, sz = C.c_void_p, C.c_size_t
P_i32 = C.POINTER(C.c_int32)
P_f64 = C.POINTER(C.c_double)
P_u8 = C.POINTER(C.c_uint8)
P_u32 = C.POINTER(C.c_uint32)
P_u64 = C.POINTER(C.c_uint64)
P_i64 = C.POINTER(C.c_int64)
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    # phylo_hip.h
    "phy_words": (i64, [i64]),
    "phy_ctx_create": (C.c_int, [C.c_int, vp, PP]),
    "phy_ctx_destroy": (None, [vp]),
    "phy_last_error": (C.c_char_p, [vp]),
    "phy_ctx_sync": (C.c_int, [vp]),
    "phy_ctx_device": (C.c_int, [vp]),
    "phy_ctx_stream": (vp, [vp]),
    "phy_dev_alloc": (C.c_int, [vp, sz, PP]),
    "phy_dev_free": (C.c_int, [vp, vp]),
    "phy_copy_h2d": (C.c_int, [vp, vp, vp, sz]),
    "phy_copy_d2h": (C.c_int, [vp, vp, vp, sz]),
    "phy_topo_create": (C.c_int, [vp, i32, P_i32, i32, PP]),
    "phy_topo_destroy": (None, [vp]),
    "phy_topo_info": (C.c_int, [vp, P_i32, P_i32, P_i32]),
    "phy_topo_programs": (C.c_int, [vp, P_i32, P_i32, P_i32, P_i32]),
    "phy_topo_split": (C.c_int, [vp, P_i32, P_i32]),
    "phy_topo_split_program": (C.c_int, [vp, C.c_int32, C.c_int32, P_i32, P_i32, P_i32, P_i32]),
    "phy_fitch_median2": (C.c_int, [vp, vp, vp, vp, i64, vp, vp]),
    "phy_fitch_median3": (C.c_int, [vp, vp, vp, vp, vp, i64, vp, vp]),
    "phy_fitch_distance1": (C.c_int, [vp, vp, vp, i64, vp, vp]),
    "phy_fitch_distance2": (C.c_int, [vp, vp, vp, vp, i64, vp, vp]),
    "phy_fitch_adjust3": (C.c_int, [vp, vp, vp, vp, vp, vp, i64, vp]),
    "phy_fitch_select": (C.c_int, [vp, vp, i64, vp, P_i64, i64, vp, vp]),
    "phy_fitch_downpass": (C.c_int, [vp, vp, vp, i64, i64, vp, vp, vp, vp]),
    "phy_fitch_uppass": (C.c_int, [vp, vp, vp, vp, i64, i64]),
    "phy_sankoff_downpass": (C.c_int, [vp, vp, i32, vp, vp, i64, i64, vp, vp, vp]),
    "phy_fitch_insertion_costs": (C.c_int, [vp, vp, vp, i64, i64, vp, i64, vp, vp]),
    "phy_diagonalize_gtr": (C.c_int, [i32, P_f64, P_f64, P_f64, P_f64]),
    "phy_diagonalize_sym": (C.c_int, [i32, P_f64, P_f64]),
    "phy_compose_gtr": (C.c_int, [i32, P_f64, P_f64, P_f64, f64, P_f64]),
    "phy_compose_sym": (C.c_int, [i32, P_f64, P_f64, f32, P_f64]),
    "phy_gamma_rates": (C.c_int, [i32, f64, f64, i32, P_f64]),
    "phy_gtr_q": (C.c_int, [i32, P_f64, P_f64, P_f64]),
    "phy_model_create": (C.c_int, [vp, i32, i32, P_f64, P_f64, P_f64, P_f64, PP]),
    "phy_model_destroy": (None, [vp]),
    "phy_model_pmatrices": (C.c_int, [vp, vp, vp, i32, vp, vp, vp]),
    "phy_lk_tree_work_size": (C.c_int, [vp, vp, i64, i32, i32, C.POINTER(sz)]),
    "phy_lk_tree": (C.c_int, [vp, vp, vp, i32, i64, vp, i64, vp, vp, vp, vp, vp, vp, vp, sz]),
    "phy_lk_tree_partials": (C.c_int, [vp, vp, i64, i32, i32, vp, i32, PP, PP, P_i64]),
    "phy_lk_tree_noalloc": (C.c_int, [vp, vp, vp, i32, i64, vp, i64, vp, vp, vp, vp]),
    "phy_ctx_kernel_timer": (C.c_int, [vp, i32]),
    "phy_ctx_kernel_time": (C.c_int, [vp, C.c_char_p, P_f64, P_i64]),
    "phy_copy_stream": (C.c_int, [vp, vp, vp, sz]),
    "phy_lk_leaf": (C.c_int, [vp, vp, i32, i64, vp, vp, vp, i64]),
    "phy_lk_median2": (C.c_int, [vp, vp, i32, i64, vp, vp, f64, vp, vp, f64, vp, vp, i64]),
    "phy_lk_root": (C.c_int, [vp, vp, i32, i64, vp, vp, i64, vp, vp]),
    "phy_lk_median3": (C.c_int, [vp, vp, i32, i64, vp, vp, f64, vp, vp, f64, vp, vp, f64, vp, vp,
                                 i64, vp, vp]),
    "phy_lk_edge": (C.c_int, [vp, vp, i32, i64, vp, vp, vp, vp, i64, f64, vp, vp]),
    "phy_lk_adjust3": (C.c_int, [vp, vp, i32, i64, vp, vp, vp, vp, i64, vp, f64, f64, f64, i32,
                                 f64, P_f64, P_f64, P_i32]),
    "phy_newton_init": (C.c_int, [vp, vp, vp]),
    "phy_newton_propose": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "phy_newton_accept": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "phy_spr_moves": (C.c_int, [i32, P_i32, i32, P_i64, P_i32, i64]),
    "phy_spr_apply": (C.c_int, [i32, P_i32, i32, P_f64, i32, i32, P_i32, P_i32, P_f64]),
    "phy_tbr_moves": (C.c_int, [i32, P_i32, i32, i32, i32, P_i64, P_i32, i64]),
    "phy_tbr_apply": (C.c_int, [i32, P_i32, i32, P_f64, i32, i32, i32, P_i32, P_i32, P_f64]),
    "phy_nni_moves": (C.c_int, [i32, P_i32, i32, P_i64, P_i32, i64]),
    "phy_lk_spr_work_size": (C.c_int, [i32, vp, i64, i32, i32, C.POINTER(sz)]),
    "phy_lk_spr_scores": (C.c_int, [vp, vp, i32, i64, vp, i64, vp, i32, P_i32, i32, P_f64, i64,
                                    P_i32, i32, vp, vp, sz]),
    "phy_lk_tbr_scores": (C.c_int, [vp, vp, i32, i64, vp, i64, vp, i32, P_i32, i32, P_f64, i64,
                                    P_i32, i32, vp, vp, sz]),
    "phy_lk_spr_incr_work_size": (C.c_int, [i32, P_i32, i32, vp, i64, i32, i32, C.POINTER(sz)]),
    "phy_lk_spr_scores_incr": (C.c_int, [vp, vp, i32, i64, vp, i64, vp, i32, P_i32, i32, P_f64,
                                         i64, P_i32, i32, vp, vp, vp, sz]),
    "phy_spr_plan_create": (C.c_int, [vp, vp, i32, i64, i32, P_i32, i32, i64, P_i32, i32, PP]),
    "phy_spr_plan_work_size": (C.c_int, [vp, C.POINTER(sz)]),
    "phy_spr_plan_run": (C.c_int, [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, sz]),
    "phy_spr_plan_destroy": (None, [vp]),
    "phy_lk_tbr_incr_work_size": (C.c_int, [i32, P_i32, i32, vp, i64, i32, i64, P_i32, i32,
                                            C.POINTER(sz)]),
    "phy_lk_tbr_scores_incr": (C.c_int, [vp, vp, i32, i64, vp, i64, vp, i32, P_i32, i32, P_f64,
                                         i64, P_i32, i32, vp, vp, vp, sz]),
    "phy_tbr_plan_create": (C.c_int, [vp, vp, i32, i64, i32, P_i32, i32, i64, P_i32, i32, PP]),
    "phy_mlp_create": (C.c_int, [vp, i32, i32, u64, PP]),
    "phy_mlp_destroy": (None, [vp]),
    "phy_mlp_n_params": (i64, [vp]),
    "phy_mlp_get_params": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    "phy_mlp_set_params": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    "phy_mlp_forward": (C.c_int, [vp, vp, vp, i64, vp]),
    "phy_mlp_backward": (C.c_int, [vp, vp, vp, vp, i64, vp, vp]),
    "phy_mlp_step": (C.c_int, [vp, vp, vp, vp, i64, f32, vp]),
    "phy_aa_batch_create": (C.c_int, [vp, i32, P_i32, P_i64, P_i32, P_i32, P_f64, P_f64, P_f64,
                                      i32, P_f64, P_f64, P_u8, P_u32, i32, PP]),
    "phy_aa_batch_destroy": (None, [vp]),
    "phy_aa_batch_info": (C.c_int, [vp, P_i64, P_i64, P_i32]),
    "phy_aa_batch_run": (C.c_int, [vp, vp, vp]),
    # phylo_synth.h
    "phy_synth_yule": (C.c_int, [i32, u64, f64, P_i32, P_i32, P_f64]),
    "phy_synth_fitch_leaves": (C.c_int, [i32, i64, u64, P_u64, i64]),
    "phy_synth_dirichlet": (C.c_int, [i32, f64, u64, P_f64]),
    "phy_synth_gtr_params": (C.c_int, [i32, f64, u64, P_f64, P_f64]),
    "phy_synth_simulate": (C.c_int, [i32, P_i32, i32, P_f64, i32, P_f64, P_f64, i32, P_f64,
                                     i64, u64, i32, P_u8]),
    "phy_synth_compress": (C.c_int, [i32, i64, P_u8, P_u8, P_u32, P_i64]),
    "phy_synth_poisson_weights": (C.c_int, [i64, f64, u64, P_u32]),
    "phy_pack_dna": (C.c_int, [i32, i64, P_u8, i64, P_u64, i64]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _fn = getattr(lib, _name, None)
    if _fn is None:
        continue  # reported by exported_symbols(); tests require every one
    _fn.restype = _res
    _fn.argtypes = _args


def exported_symbols() -> dict[str, bool]:
    """Which declared symbols the loaded library exports."""
    return {n: hasattr(lib, n) for n in SIGNATURES}


class PhyError(RuntimeError):
    """A non-zero status from the C ABI (the OCaml stubs raise Failure)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"phylo_hip error {code}: {msg}")
        self.code = code


STATUS = {0: "PHY_OK", 1: "PHY_EINVAL", 2: "PHY_EDEVICE", 3: "PHY_ENOMEM",
          4: "PHY_ENUMERIC", 5: "PHY_ERANGE"}

PHY_F64, PHY_F32, PHY_FUSED = 0, 1, 2
PHY_GAMMA_REF, PHY_GAMMA_MEAN = 0, 1


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != 0:
        msg = ""
        if ctx is not None:
            m = lib.phy_last_error(ctx)
            msg = m.decode() if m else ""
        raise PhyError(rc, f"{what}: {STATUS.get(rc, rc)} {msg}".strip())


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array."""
    return a.ctypes.data_as(C.POINTER(ctype))
