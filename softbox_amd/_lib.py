"""ctypes binding of libsoftbox_hip.so (the C ABI declared in include/softbox_hip.h).

The library is the product path: if it is missing or cannot be loaded this module raises,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsoftbox_hip.so")

SBX_OK = 0
SBX_NOALLOC_MAX_KEYS = 4194304     # sbx_csr_build_noalloc limits (include/softbox_hip.h)
SBX_NOALLOC_MAX_ITEMS = 2097152
SBX_ERR_ARG = -1
SBX_ERR_HIP = -2
SBX_ERR_NOMEM = -3
SBX_ERR_STATE = -4
SBX_RNG_WANG = 0
SBX_RNG_PHILOX = 1
SBX_MAX_CAND = 256
SBX_XL_STRIDE = 12
SBX_STRAIN_WIN = 10
SBX_TORQUE_WIN = 5

vp = C.c_void_p
i32 = C.c_int
u32 = C.c_uint32
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int)


class SbxFilaments(C.Structure):
    _fields_ = [("n", i32)] + [
        (name, vp)
        for name in (
            "coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam",
            "brownTransScale", "brownRotScale", "forceSum", "torqueSum",
            "end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide",
        )
    ]


class SbxCrosslinks(C.Structure):
    _fields_ = [("nLinks", i32)] + [
        (name, vp)
        for name in (
            "linkFilA", "linkFilB", "loc1", "loc2", "linkState", "strainHist", "strainPlace",
            "linkOrientSame", "torqueMagHist", "torqueMagPlace", "xlinkData", "activeLinkCount",
        )
    ]


class SbxNetParams(C.Structure):
    _fields_ = [
        ("params", C.c_float * 2),
        ("chainParams", C.c_float * 7),
        ("chainOn", i32),
        ("xlParams", C.c_float * 3),
        ("offParams", C.c_float * 5),
        ("torsionParams", C.c_float * 2),
        ("unbindOn", i32),
        ("boxParams", C.c_float * 8),
        ("confineOn", i32),
        ("gridOn", i32),
        ("periodic", i32),
        ("maxCand", i32),
        ("gridParams", C.c_float * 6),
        ("gridDims", C.c_int * 4),
        ("boxLen", C.c_float * 3),
        ("actinRadius", C.c_float),
        ("seed", i32),
        ("rng", i32),
        ("stream", u32),
        ("formOn", i32),
        ("formInt", i32),
        ("formParams", C.c_float * 7),
        ("formMode", i32),
        ("formGridParams", C.c_float * 6),
        ("formGridDims", C.c_int * 4),
        ("formReqCap", i32),
    ]


class SbxMotors(C.Structure):
    _fields_ = (
        [("nM", i32)]
        + [(k, vp) for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam",
                             "brownTransScale", "brownRotScale", "forceSum", "torqueSum", "head", "hUVec", "rodUVec",
                             "reach", "boundSeg", "cooldown", "stats", "capStats", "nucleotideState",
                             "forceDotPlace", "bindArc", "forceDotFil", "forceMag", "forceDotHist", "bondData",
                             "reachSeg", "reachCount")]
        + [("nD", i32)]
        + [(k, vp) for k in ("motorA", "motorB", "parallel")]
        + [("nBb", i32), ("nMD", i32)]
        + [(k, vp) for k in ("bbCoord", "bbUVec", "bbYVec", "bbZVec", "bbEnd1", "bbEnd2", "bbSegLength",
                             "bbTransGam", "bbRotGam", "bbBrownTransScale", "bbBrownRotScale", "bbInvDragY",
                             "bbForceSum", "bbTorqueSum", "headBackboneSlot", "tetherMotor", "attachAxial",
                             "miniData", "radial", "attachCoeffK")]
    )


class SbxMotorParams(C.Structure):
    _fields_ = [
        ("bodyParams", C.c_float * 2),
        ("jointParams", C.c_float * 11),
        ("kinParams", C.c_float * 14),
        ("nucParams", C.c_float * 8),
        ("xbParams", C.c_float * 6),
        ("dimerParams", C.c_float * 5),
        ("miniParams", C.c_float * 3),
        ("bbBodyParams", C.c_float * 2),
        ("seedMotor", i32),
        ("seedBackbone", i32),
        ("nodeMode", i32),
    ]


SBX_SKIN_REBUILD = 0
SBX_SKIN_FIXED = 1
SBX_SKIN_LEARNED = 2


class SbxSkinParams(C.Structure):
    _fields_ = [("skinDefault", C.c_float), ("skinMin", C.c_float), ("skinMax", C.c_float), ("margin", C.c_float),
                ("listCap", i32), ("horizon", i32), ("extraTrigger", i32), ("extraCap", i32)]


class SbxReachSkinParams(C.Structure):
    _fields_ = [("skinDefault", C.c_float), ("skinMin", C.c_float), ("skinMax", C.c_float), ("margin", C.c_float),
                ("cellFloor", C.c_float), ("horizon", i32), ("extraTrigger", i32), ("extraCap", i32)]


class SbxTurnover(C.Structure):
    _fields_ = [("C", i32)] + [(name, vp) for name in (
        "filState", "monomerCount", "seedNode", "nucFrac", "cofFrac", "bTransDiff", "bRotDiff", "pool", "ledger",
        "stats")]


class SbxTurnoverParams(C.Structure):
    _fields_ = [("interval", i32), ("seed", i32), ("growthOn", i32), ("proxy", i32),
                ("growParams", C.c_float * 3), ("splitParams", C.c_float * 3), ("dragParams", C.c_float * 9),
                ("depolyParams", C.c_float * 3), ("depolyRateParams", C.c_float * 2), ("agingParams", C.c_float * 2),
                ("cofilinParams", C.c_float * 2), ("birthParams", C.c_float * 2),
                ("onRate", C.c_double), ("biochemDeltaT", C.c_double), ("uMPerMonomer", C.c_double)]


# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "sbx_create": [C.POINTER(vp), i32, vp],
    "sbx_destroy": [vp],
    "sbx_last_error": [vp],
    "sbx_clear_error": [vp],
    "sbx_sync": [vp],
    "sbx_alloc": [vp, C.POINTER(vp), C.c_size_t],
    "sbx_free": [vp, vp],
    "sbx_push": [vp, vp, vp, C.c_size_t],
    "sbx_pull": [vp, vp, vp, C.c_size_t],
    "sbx_copy": [vp, vp, vp, C.c_size_t],
    "sbx_fill_i32": [vp, vp, i32, C.c_size_t],
    "sbx_version": [],
    "sbx_zero_accumulators": [vp, vp, vp, i32],
    "sbx_brownian_force": [vp, vp, vp, vp, vp, vp, vp, f32p, i32p, i32, i32, u32],
    "sbx_integrate": [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, f32p, i32p, i32],
    "sbx_derive": [vp, vp, vp, vp, vp, vp, vp, vp, i32p, i32],
    "sbx_force_build": [vp, vp, vp, vp, vp, vp, vp, vp, vp, f32p, i32p, i32, i32, u32],
    "sbx_integ_derive": [vp] + [vp] * 13 + [f32p, i32],
    "sbx_integ_derive_confined": [vp] + [vp] * 13 + [f32p, i32p, i32],
    "sbx_confine": [vp, vp, vp, vp, vp, vp, vp, f32p, i32p, i32],
    "sbx_chain_forces": [vp] + [vp] * 11 + [f32p, i32p, i32],
    "sbx_publish_filaments": [vp, vp, vp, vp, vp, vp, vp, f32p, i32, i32],
    "sbx_body_cell": [vp, vp, f32p, i32p, i32p, vp, i32, i32],
    "sbx_grid_build": [vp, vp, i32p, i32p, vp, vp, i32],
    "sbx_broad_phase": [vp, vp, vp, vp, vp, vp, i32p, f32p, i32p, vp, vp, i32, i32, i32],
    "sbx_csr_build": [vp, vp, i32, i32, vp, vp],
    "sbx_csr_build_noalloc": [vp, vp, i32, i32, vp, vp],
    "sbx_count_active_links": [vp, vp, vp, vp, vp, i32, i32],
    "sbx_filid": [vp, vp, vp, vp, vp, i32],
    "sbx_grid_form_candidates": [vp] + [vp] * 3 + [f32p, i32p] + [vp] * 6 + [f32p, vp, i32, i32, i32, f32p],
    "sbx_form_gates": [vp] + [vp] * 11 + [f32p, i32p, i32, i32, i32, u32, i32, f32p],
    "sbx_form_admit_reduce": [vp, vp, vp, vp, vp, i32, i32],
    "sbx_form_admit": [vp] + [vp] * 13 + [f32p, i32, i32, i32],
    "sbx_form_allocate": [vp] + [vp] * 20 + [i32, i32],
    "sbx_net_form_stats": [vp, C.POINTER(C.c_longlong), i32],
    "sbx_aging_age": [vp, vp, vp, f32p, i32, i32],
    "sbx_aging_growth_atp": [vp, vp, vp, vp, i32],
    "sbx_aging_split_inherit_nuc": [vp, vp, vp, vp, vp, i32, i32],
    "sbx_depoly": [vp] + [vp] * 8 + [f32p, f32p, i32, i32, i32, i32, i32, u32],
    "sbx_apply_death": [vp] + [vp] * 10 + [i32],
    "sbx_growth_grow": [vp] + [vp] * 5 + [f32p, i32, i32, i32, i32, i32, u32],
    "sbx_growth_mark_splits": [vp] + [vp] * 9 + [f32p, i32, i32],
    "sbx_growth_split_wire": [vp] + [vp] * 11 + [f32p, i32, i32],
    "sbx_growth_recompute_drag": [vp] + [vp] * 8 + [f32p, i32, i32],
    "sbx_cofilin_accumulate": [vp, vp, vp, vp, f32p, i32, i32],
    "sbx_cofilin_dissolve": [vp] + [vp] * 5 + [f32p, i32, i32],
    "sbx_cofilin_fresh": [vp, vp, vp, vp, vp, i32, i32],
    "sbx_birth_alloc": [vp] + [vp] * 14 + [f32p, i32, i32, i32],
    "sbx_turnover_cadence": [vp, C.POINTER(SbxFilaments), C.POINTER(SbxTurnover), C.POINTER(SbxTurnoverParams), i32,
                             i32, u32],
    "sbx_net_set_turnover": [vp, C.POINTER(SbxTurnover), C.POINTER(SbxTurnoverParams)],
    "sbx_link_unbind": [vp] + [vp] * 10 + [f32p, i32p, i32, i32, u32],
    "sbx_link_forces": [vp] + [vp] * 10 + [f32p, i32, i32],
    "sbx_link_torsion": [vp] + [vp] * 8 + [f32p, i32, i32],
    "sbx_seg_gather": [vp, vp, vp, vp, vp, vp, vp, i32, i32],
    "sbx_publish_head_from_body": [vp, vp, vp, vp, vp, vp, vp, i32p, i32],
    "sbx_publish_motors": [vp, vp, vp, vp, vp, vp, vp, i32p, i32p, i32, i32],
    "sbx_grid_reachable": [vp] + [vp] * 5 + [f32p, i32p] + [vp] * 6 + [f32p, i32p, i32, i32],
    "sbx_bind_nearest": [vp] + [vp] * 9 + [f32p, i32p, i32, i32],
    "sbx_catch_slip_release": [vp] + [vp] * 6 + [f32p, i32p, i32, i32, u32],
    "sbx_nucleotide_cycle": [vp, vp, vp, vp, f32p, i32p, i32, i32, u32],
    "sbx_bond_forces": [vp] + [vp] * 13 + [f32p, i32, i32],
    "sbx_apply_head_force": [vp, vp, vp, vp, i32],
    "sbx_register_force_dot": [vp] + [vp] * 6 + [i32],
    "sbx_motor_seg_gather": [vp] + [vp] * 5 + [i32],
    "sbx_motor_joints": [vp] + [vp] * 8 + [f32p, i32],
    "sbx_dimer_couple": [vp] + [vp] * 10 + [f32p, vp, i32, i32],
    "sbx_mini_tether": [vp] + [vp] * 14 + [f32p, i32, i32, i32],
    "sbx_backbone_gather": [vp] + [vp] * 5 + [i32],
    "sbx_node_tether": [vp] + [vp] * 14 + [f32p, i32, i32, i32],
    "sbx_node_tether_bound_stiffen": [vp] + [vp] * 15 + [f32p, i32, i32, i32],
    "sbx_node_xbridge_inject": [vp] + [vp] * 8 + [i32, i32],
    "sbx_bind_nearest_node_aware": [vp] + [vp] * 10 + [f32p, i32p, i32, i32],
    "sbx_brute_reachable": [vp] + [vp] * 7 + [f32p, i32p, i32, i32],
    "sbx_brute_reachable_node_aware": [vp] + [vp] * 8 + [f32p, i32p, i32, i32],
    "sbx_nuc_count_bound_fil": [vp, vp, vp, i32, i32],
    "sbx_nuc_emit": [vp] + [vp] * 6 + [f32p, i32, i32, i32, i32, i32, i32, u32],
    "sbx_nuc_tag_seeds": [vp] + [vp] * 4 + [i32, i32],
    "sbx_nuc_init_newborn": [vp] + [vp] * 5 + [f32p, i32, i32],
    "sbx_nuc_seed_tether": [vp] + [vp] * 9 + [f32p, i32, i32],
    "sbx_nuc_seed_tether_node_react": [vp] + [vp] * 8 + [f32p, i32, i32],
    "sbx_nuc_seed_aim_torque": [vp] + [vp] * 4 + [f32p, i32],
    "sbx_nuc_dissolve": [vp, vp, f32p, i32, i32, i32, i32, u32],
    "sbx_net_create": [vp, C.POINTER(SbxFilaments), C.POINTER(SbxCrosslinks), C.POINTER(SbxNetParams), C.POINTER(vp)],
    "sbx_net_destroy": [vp],
    "sbx_net_attach_motors": [vp, C.POINTER(SbxMotors), C.POINTER(SbxMotorParams)],
    "sbx_net_step": [vp, i32, i32],
    "sbx_net_publish": [vp],
    "sbx_net_topology_changed": [vp],
    "sbx_net_candidates": [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)],
    "sbx_net_set_profiling": [vp, i32],
    "sbx_net_kernel_times": [vp, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.POINTER(C.c_longlong), i32],
    "sbx_net_time_broad_phase": [vp, i32, C.POINTER(C.c_float)],
    "sbx_net_use_graph": [vp, i32],
    "sbx_net_set_overlap": [vp, i32],
    "sbx_net_set_shape": [vp, i32, i32, i32, i32],
    "sbx_net_set_motor_group": [vp, i32],
    "sbx_capture_check": [vp, i32],
    "sbx_net_capture_log": [vp, vp, i32],
    "sbx_net_motor_group": [vp],
    "sbx_net_tune": [vp, i32, i32, C.POINTER(i32), C.POINTER(i32), f32p],
    "sbx_batch_create": [vp, i32, C.POINTER(SbxFilaments), C.POINTER(SbxCrosslinks), C.POINTER(SbxNetParams),
                         C.POINTER(vp)],
    "sbx_batch_destroy": [vp],
    "sbx_batch_step": [vp, i32, i32],
    "sbx_batch_publish": [vp],
    "sbx_batch_net": [vp, C.POINTER(vp)],
    "sbx_batch_layout": [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)],
    "sbx_net_set_skin": [vp, i32, C.POINTER(SbxSkinParams), vp],
    "sbx_net_skin_stats": [vp, C.POINTER(C.c_longlong), i32],
    "sbx_net_set_reach_skin": [vp, i32, C.POINTER(SbxReachSkinParams), vp],
    "sbx_net_reach_skin_stats": [vp, C.POINTER(C.c_longlong), i32],
    "sbx_net_skin_state": [vp] + [C.POINTER(vp)] * 5 + [C.POINTER(i32), f32p],
    "sbx_skin_features": [vp, vp, vp, i32, i32, vp, vp, vp, vp, vp, C.c_float, C.c_float, i32, f32p, vp],
    "sbx_skin_track": [vp, vp, vp, i32, i32, f32p, vp, vp, vp, i32],
    "sbx_skin_cell_reduce": [vp, vp, vp, i32, vp, vp, i32, C.c_float, vp],
    "sbx_skin_cell_scatter": [vp, vp, vp, i32, vp, vp],
    "sbx_mlp_nparam": [],
    "sbx_mlp_forward": [vp, vp, vp, i32, vp],
    "sbx_mlp_scratch_floats": [i32],
    "sbx_mlp_loss_grad": [vp, vp, vp, vp, i32, f32p, vp, vp, vp],
    "sbx_mlp_adam": [vp, vp, vp, vp, vp, i32, C.c_float, C.c_float, C.c_float, C.c_float],
}
_RESTYPES = {"sbx_destroy": None, "sbx_last_error": C.c_char_p, "sbx_clear_error": None,
             "sbx_version": C.c_char_p, "sbx_net_destroy": None, "sbx_batch_destroy": None, "sbx_mlp_scratch_floats": C.c_size_t}

EXPORTED = tuple(_SIGS)

_lib = None


class SoftboxError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"softbox_hip error {code}: {message}")
        self.code = code


def load() -> C.CDLL:
    """Load libsoftbox_hip.so (built in-tree); raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python -m softbox_amd.build` (the HIP library has no fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, i32)
    _lib = lib
    return lib
