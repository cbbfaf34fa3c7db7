"""ctypes prototypes for include/cmf_host.h and include/cmf_hip.h."""
from __future__ import annotations

import ctypes as C

P = C.c_void_p
I = C.c_int
I32P = C.POINTER(C.c_int32)
DP = C.POINTER(C.c_double)

class cmfh_channel(C.Structure):
    _fields_ = [("n", C.c_int32), ("nlfp", C.c_int32), ("lfldout", C.c_int32), ("next", P)] + \
               [(k, P) for k in ("elevtn rivelv rivhgt rivwth rivlen rivman nxtdst grarea dwnelv "
                                 "rivstomax fldstomax fldgrd").split()] + \
               [(k, C.c_double) for k in ("dfrcinc", "pmanfld", "pdstmth")]


HOST_FUNCS = {
    "cmfh_calc_1d_seq": (I, [I, I, P, P, P, P, P, P]),
    "cmfh_build_upstream_csr": (I, [I, I, P, P, P]),
    "cmfh_set_fldstg": (I, [I, I, I, P, P, P, P, P, P, P, P, P, P, P]),
    "cmfh_levee_init": (I, [I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "cmfh_storage_sea_surface": (I, [I, I, P, P, P, P, P, P, P, P, P]),
    "cmfh_priority_flood": (I, [I, I, P, P, C.c_float, P, P, P]),
    "cmfh_upstream_area": (I, [I, I, P, P, P, P]),
    "cmfh_accumulate": (I, [I, P, I, P, P]),
    "cmfh_steady_state": (I, [C.POINTER(cmfh_channel), P, P, P, P, P, P]),
    "cmfh_accumulate_gain": (I, [I, P, I, P, P, P, P, P, P]),
    "cmfh_periodic_response": (I, [C.POINTER(cmfh_channel), P, I, P, P, P, P, P, P, P]),
    "cmfh_storage_at_level": (I, [C.POINTER(cmfh_channel), P, P, P]),
}


class cmf_dims(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nseqall", "nseqriv", "nlfp", "npthout", "npthlev", "ndam")]


class cmf_options(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("ladpstp", "lfplain", "lfldout", "lpthout", "ldamout",
                                           "llevee", "precision")] + \
               [(k, C.c_double) for k in ("pgrv", "pmanfld", "pdstmth", "pcadp", "pminslp")] + \
               [("jprb", C.c_int32), ("ldamh22", C.c_int32)]


MAP_FIELDS = [("i1next", P)] + [(k, P) for k in (
    "grarea elevtn nxtdst rivlen rivwth rivhgt rivman rivstomax rivelv dwnelv "
    "fldhgt fldstomax fldgrd").split()] + [("dfrcinc", C.c_double)] + [(k, P) for k in (
    "i2mask pth_upst pth_down pth_dst pth_elv pth_wth pth_man levhgt levfrc bashgt levdst "
    "levbassto levtopsto levfilsto i1dam damseq damstat dam_convol dam_emevol dam_adjvol "
    "dam_qa dam_qn dam_qf dam_norvol dam_r_volupa").split()]


class cmf_map(C.Structure):
    _fields_ = MAP_FIELDS


STATE_FIELDS = ("rivsto fldsto levsto damsto daminf rivout fldout rivout_pre fldout_pre "
                "rivdph_pre fldsto_pre pthflw pthflw_pre rivdph flddph fldfrc fldare sfcelv storge "
                "outflw rivvel rivinf fldinf pthout levdph pthflwsum rivout_avg fldout_avg outflw_avg "
                "rivvel_avg pthout_avg runoff_avg outflw_max rivdph_max storge_max daminf_avg "
                "pthflw_avg pthflwsum_avg").split()


class cmf_state(C.Structure):
    _fields_ = [(k, P) for k in STATE_FIELDS]


OUTPUT_FIELDS = ("rivout fldout outflw rivvel pthout runoff daminf outflw_max rivdph_max storge_max "
                 "pthflw").split()


class cmf_output(C.Structure):
    _fields_ = [(k, P) for k in OUTPUT_FIELDS]


class cmf_ensemble(C.Structure):
    _fields_ = [("ndom", C.c_int32), ("dom_cells", P), ("members", C.c_int32)]


TUNE_FIELDS = ("bs_outflow", "bs_rate", "bs_update", "fused", "bs_fused", "grid_fused", "xcd", "graph")


class cmf_tune(C.Structure):
    _fields_ = [(k, C.c_int32) for k in TUNE_FIELDS]


H = C.c_void_p
HIP_FUNCS = {
    "cmf_hip_create": (I, [C.POINTER(cmf_dims), C.POINTER(cmf_options), C.POINTER(H)]),
    "cmf_hip_create_ensemble": (I, [C.POINTER(cmf_dims), C.POINTER(cmf_options), C.POINTER(cmf_ensemble),
                                    C.POINTER(H)]),
    "cmf_hip_set_runoff_scaled": (I, [H, P, P]),
    "cmf_hip_set_inpmat": (I, [H, C.c_int32, C.c_int32, C.c_int32, P, P, P, C.c_double, C.c_double]),
    "cmf_hip_set_runoff_grid": (I, [H, P]),
    "cmf_hip_interp_runoff": (I, [H, P]),
    "cmf_hip_set_runoff_records": (I, [H, C.c_int32, P]),
    "cmf_hip_use_runoff_record": (I, [H, C.c_int32]),
    "cmf_hip_set_dwnelv": (I, [H, P]),
    "cmf_hip_set_upsinf": (I, [H, P]),
    "cmf_hip_set_kinematic": (I, [H, C.c_int32]),
    "cmf_hip_set_slopemix": (I, [H, C.c_int32]),
    "cmf_hip_restart_write": (I, [H, C.c_char_p, C.c_int32, C.c_int32, P, P, C.c_int32, C.c_int32, C.c_double,
                                  C.c_double]),
    "cmf_hip_restart_read": (I, [H, C.c_char_p, C.c_int32, C.c_int32, P, P, C.c_int32, C.c_int32]),
    "cmf_hip_output_write": (I, [H, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, P, P, C.c_int32, C.c_double]),
    "cmf_hip_update_boundary": (I, [H, P, P]),
    "cmf_hip_diag_reset_output": (I, [H]),
    "cmf_hip_diag_avemax_output": (I, [H, C.c_double]),
    "cmf_hip_diag_getave_output": (I, [H, C.POINTER(C.c_double)]),
    "cmf_hip_get_output": (I, [H, C.POINTER(cmf_output)]),
    "cmf_hip_ensemble_advance": (I, [H, C.c_double, C.c_int32, P, P]),
    "cmf_hip_instance_times": (I, [H, P]),
    "cmf_hip_get_dt_min": (I, [H, P]),
    "cmf_hip_instance_storage": (I, [H, P]),
    "cmf_hip_set_tune": (I, [H, C.POINTER(cmf_tune)]),
    "cmf_hip_get_tune": (I, [H, C.POINTER(cmf_tune)]),
    "cmf_hip_autotune": (I, [H, C.c_double, C.c_int32, C.POINTER(cmf_tune), P, C.c_int32]),
    "cmf_hip_destroy": (None, [H]),
    "cmf_hip_last_error": (C.c_char_p, []),
    "cmf_hip_set_map": (I, [H, C.POINTER(cmf_map)]),
    "cmf_hip_set_state": (I, [H, C.POINTER(cmf_state)]),
    "cmf_hip_set_state_shared": (I, [H, C.POINTER(cmf_state)]),
    "cmf_hip_get_state": (I, [H, C.POINTER(cmf_state)]),
    "cmf_hip_set_runoff": (I, [H, P]),
    "cmf_hip_physics_fldstg": (I, [H]),
    "cmf_hip_calc_outpre": (I, [H]),
    "cmf_hip_physics_advance": (I, [H, C.POINTER(C.c_double), C.POINTER(C.c_int32), P, C.c_int32]),
    "cmf_hip_physics_advance_capped": (I, [H, C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32),
                                           P, C.c_int32]),
    "cmf_hip_physics_advance_noalloc": (I, [H, C.POINTER(C.c_double), C.POINTER(C.c_int32), P, C.c_int32]),
    "cmf_hip_synchronize": (I, [H]),
    "cmf_hip_kernel_times": (I, [H, C.c_int32, P]),
}


MLP_FUNCS = {
    "cmf_mlp_create": (I, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(H)]),
    "cmf_mlp_destroy": (None, [H]),
    "cmf_mlp_last_error": (C.c_char_p, []),
    "cmf_mlp_param_count": (C.c_int64, [H]),
    "cmf_mlp_set_loss": (I, [H, C.c_float]),
    "cmf_mlp_set_params": (I, [H, P]),
    "cmf_mlp_get_params": (I, [H, P, P]),
    "cmf_mlp_forward": (I, [H, P, C.c_int64, P]),
    "cmf_mlp_grad": (I, [H, P, P, C.c_int32, P]),
    "cmf_mlp_set_data": (I, [H, P, P, C.c_int64]),
    "cmf_mlp_train": (I, [H, P, C.c_int64, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float,
                          C.c_int32, P]),
    "cmf_mlp_gemm_test": (I, [C.c_int32, C.c_int32, C.c_int32, P, C.c_int32, C.c_int32, P, C.c_int32,
                              C.c_int32, P]),
}


def _declare(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def declare_host(lib):
    _declare(lib, HOST_FUNCS)


def declare_hip(lib):
    _declare(lib, HIP_FUNCS)
    _declare(lib, MLP_FUNCS)


def ptr(a):
    """Raw data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data
