"""ctypes mirror of include/geauxdock_amd.h (structs, enums, prototypes).

The product is the C-ABI library ``libgeauxdock_amd.so`` (C++ host code plus
hand-written HIP kernels for gfx950).  This module only declares its types so
Python callers (tests, bench, smoke) can reach it; it computes nothing.
"""
import ctypes as C

GD_NPT, GD_NLT, GD_NRES, GD_NELE, GD_NTERM = 31, 24, 20, 50, 10
GD_MCS_INVALID = 0.01234
GD_BADTYPE = 50
TERMS = ("vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total")

GD_OK, GD_ERR_ARG, GD_ERR_HIP, GD_ERR_SIZE, GD_ERR_TYPE, GD_ERR_IO = 0, 1, 2, 3, 4, 5
GD_ERR_NOMEM, GD_ERR_STATE, GD_ERR_BUSY, GD_ERR_NODEV = 6, 7, 8, 9
STATUS_NAMES = {0: "GD_OK", 1: "GD_ERR_ARG", 2: "GD_ERR_HIP", 3: "GD_ERR_SIZE", 4: "GD_ERR_TYPE",
                5: "GD_ERR_IO", 6: "GD_ERR_NOMEM", 7: "GD_ERR_STATE", 8: "GD_ERR_BUSY",
                9: "GD_ERR_NODEV"}
GD_MOVE_REF_UNIFORM, GD_MOVE_GAUSSIAN = 0, 1
GD_EXCH_TEMP, GD_EXCH_LIG, GD_EXCH_PRT = 1, 2, 4

_f = C.c_float
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class gd_enepara(C.Structure):
    _fields_ = [
        ("p1a", _f * GD_NPT * GD_NLT), ("p2a", _f * GD_NPT * GD_NLT),
        ("lj0", _f), ("lj1", _f), ("el1", _f), ("el0", _f), ("a1", _f), ("b1", _f),
        ("pmf0", _f * GD_NPT * GD_NLT), ("pmf1", _f * GD_NPT * GD_NLT),
        ("hdb0", _f * GD_NPT * GD_NLT), ("hdb1", _f * GD_NPT * GD_NLT),
        ("hpp", _f * GD_NRES), ("hpl0", _f * GD_NLT), ("hpl1", _f * GD_NLT), ("hpl2", _f * GD_NLT),
        ("kde2", _f), ("kde3", _f),
        ("a_para", _f * GD_NTERM), ("b_para", _f * GD_NTERM), ("w", _f * GD_NTERM),
        ("ele", _f * GD_NELE),
    ]


class gd_job_desc(C.Structure):
    _fields_ = [
        ("n_lig", C.c_int), ("n_prt", C.c_int), ("n_tmp", C.c_int),
        ("lig_natom", C.c_int), ("prt_npoint", C.c_int), ("kde_npoint", C.c_int),
        ("mcs_nrow", C.c_int),
        ("lig_x", _fp), ("lig_y", _fp), ("lig_z", _fp), ("lig_t", _ip), ("lig_c", _fp),
        ("lig_perm", _ip), ("lig_center", _f * 3),
        ("prt_x", _fp), ("prt_y", _fp), ("prt_z", _fp),
        ("prt_t", _ip), ("prt_c", _ip), ("prt_cdc", _ip), ("prt_seq3r", _ip),
        ("prt_hpp", _fp), ("prt_ele", _fp), ("pocket_center", _f * 3),
        ("psp", _fp), ("psp_ncol", C.c_int),
        ("kde_t", _ip), ("kde_x", _fp), ("kde_y", _fp), ("kde_z", _fp),
        ("mcs_x", _fp), ("mcs_y", _fp), ("mcs_z", _fp), ("mcs_tcc", _fp),
        ("ene", C.POINTER(gd_enepara)), ("minus_beta", _fp),
    ]


class gd_mc_param(C.Structure):
    _fields_ = [
        ("steps_total", C.c_int), ("steps_per_dump", C.c_int), ("steps_per_exchange", C.c_int),
        ("move_mode", C.c_int), ("move_scale", _f * 6), ("seed", C.c_uint64),
        ("job_id", C.c_uint32), ("record", C.c_int), ("exchange_axes", C.c_int),
    ]


class gd_replica_out(C.Structure):
    _fields_ = [
        ("movematrix", _fp), ("energy", _fp), ("idx_tmp", _ip), ("n_accept", _ip),
        ("n_exchange", _ip), ("best_energy", _fp), ("is_accept", _ip),
        ("n_swap_lig", _ip), ("n_swap_prt", _ip),
    ]


class gd_stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double), ("ms_pair_kernel", C.c_double),
        ("ms_finish_kernel", C.c_double), ("n_pair_launch", C.c_int),
        ("pair_evals", C.c_double), ("replica_steps", C.c_double),
        ("ms_step_period", C.c_double), ("ms_pair_span", C.c_double),
    ]


class gd_tiling(C.Structure):
    _fields_ = [("atoms_per_lane", C.c_int), ("tile_points", C.c_int), ("waves_per_cu", C.c_int)]


class gd_server_config(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("sweep_steps", C.c_int),
                ("converge_sweeps", C.c_int), ("converge_tol", C.c_float)]


class gd_request_files(C.Structure):
    _fields_ = [
        ("sdf", C.c_char_p), ("ff", C.c_char_p), ("pdb", C.c_char_p), ("ligand_id", C.c_char_p),
        ("ene", C.POINTER(gd_enepara)), ("n_tmp", C.c_int), ("floor_temp", _f), ("ceiling_temp", _f),
    ]


class gd_job_result(C.Structure):
    _fields_ = [
        ("ticket", C.c_int64), ("status", C.c_int), ("t_submit", C.c_double),
        ("t_admit", C.c_double), ("t_done", C.c_double), ("replica_steps", C.c_double),
    ]


# gd_record_fn: (user, dump, step_begin, n_rep, cap, rec, count)
RECORD_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float),
                        C.POINTER(C.c_int))
GD_STREAM_BLOCKING = 1
REC_W = 17   # floats per accept record: step, movematrix[6], energy[10]

# Every extern "C" entry point of include/geauxdock_amd.h, with its ctypes
# prototype (restype, argtypes).  Tests check the library exports all of them.
_vp = C.c_void_p
PROTOTYPES = {
    "gd_init": (C.c_int, [C.c_int]),
    "gd_device_info": (C.c_int, [C.c_char_p, C.c_int, _ip, _ip, C.POINTER(C.c_size_t)]),
    "gd_last_error": (C.c_char_p, []),
    "gd_version": (C.c_char_p, []),
    "gd_set_temperature": (C.c_int, [C.c_int, _f, _f, _fp]),
    "gd_job_create": (C.c_int, [C.POINTER(gd_job_desc), C.POINTER(_vp)]),
    "gd_job_destroy": (C.c_int, [_vp]),
    "gd_job_sizes": (C.c_int, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "gd_job_device_bytes": (C.c_size_t, [_vp]),
    "gd_energy_eval": (C.c_int, [_vp, C.c_int, _ip, _fp, _fp, _fp, _fp]),
    "gd_energy_eval_noalloc": (C.c_int, [_vp, C.c_int, _ip, _fp, _fp, _fp]),
    "gd_dock": (C.c_int, [_vp, C.POINTER(gd_mc_param), C.POINTER(gd_replica_out),
                          C.POINTER(gd_stats)]),
    "gd_dock_records": (C.c_int, [_vp, C.c_int, _fp, _ip]),
    "gd_dock_stream": (C.c_int, [_vp, C.POINTER(gd_mc_param), _vp, _vp, C.c_int,
                                 C.POINTER(gd_replica_out), C.POINTER(gd_stats)]),
    "gd_record_csv_open": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "gd_record_csv_write": (None, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _ip]),
    "gd_record_csv_close": (C.c_int, [_vp]),
    "gd_job_replicas": (C.c_int, [_vp, C.POINTER(gd_replica_out)]),
    "gd_batch_dock": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(gd_mc_param),
                                C.POINTER(gd_replica_out), C.POINTER(gd_stats)]),
    "gd_tiling_get": (C.c_int, [_vp, C.POINTER(gd_tiling)]),
    "gd_job_set_streams": (C.c_int, [_vp, C.c_int]),
    "gd_tiling_set": (C.c_int, [_vp, C.POINTER(gd_tiling)]),
    "gd_autotune": (C.c_int, [_vp, C.c_int, C.POINTER(gd_tiling), C.POINTER(C.c_double)]),
    "gd_autotune_shared": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.POINTER(gd_tiling),
                                     C.POINTER(C.c_double)]),
    "gd_tiling_lookup": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(gd_tiling)]),
    "gd_server_create": (C.c_int, [C.POINTER(gd_server_config), C.POINTER(_vp)]),
    "gd_server_destroy": (C.c_int, [_vp]),
    "gd_server_submit": (C.c_int, [_vp, C.POINTER(gd_job_desc), C.POINTER(gd_mc_param),
                                   C.POINTER(gd_replica_out), C.POINTER(C.c_int64)]),
    "gd_job_footprint": (C.c_size_t, [C.POINTER(gd_job_desc)]),
    "gd_server_usage": (C.c_int, [_vp, C.POINTER(C.c_size_t), _ip]),
    "gd_server_step": (C.c_int, [_vp, _ip, _ip]),
    "gd_server_cancel": (C.c_int, [_vp, C.c_int64]),
    "gd_server_results": (C.c_int, [_vp, C.POINTER(gd_job_result), C.c_int, _ip]),
    "gd_server_clock": (C.c_double, [_vp]),
    "gd_server_submit_files": (C.c_int, [_vp, C.POINTER(gd_request_files), C.POINTER(gd_mc_param),
                                         C.POINTER(C.c_int64)]),
    "gd_server_take": (C.c_int, [_vp, C.c_int64, _ip, C.POINTER(gd_replica_out)]),
    "gd_load_enepara": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                  C.POINTER(gd_enepara)]),
    "gd_load_complex": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                  C.POINTER(gd_enepara), C.c_int, _f, _f, C.POINTER(_vp)]),
    "gd_synth_complex": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f, _f,
                                   C.c_uint64, C.POINTER(gd_enepara), C.POINTER(_vp)]),
    "gd_synth_screen": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                  _f, _f, C.POINTER(gd_enepara), C.POINTER(_vp)]),
    "gd_host_job_desc": (C.c_int, [_vp, C.POINTER(gd_job_desc)]),
    "gd_host_job_free": (C.c_int, [_vp]),
    "gd_kabsch_rmsd": (_f, [C.c_int, _fp, _fp]),
    "gd_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_int,
                         C.POINTER(C.c_uint32)]),
    "gd_move_draws": (C.c_int, [C.c_uint64, C.c_uint32, C.c_int, _fp, C.c_int, C.c_uint32,
                                C.c_int, C.c_int, _fp]),
    "gd_move_incr_words": (C.c_int, [C.c_int, _fp, C.c_int, C.POINTER(C.c_uint32), _fp]),
    "gd_measure_sclk": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "gd_device_copy": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "gd_far_r2": (C.c_float, [C.c_float, C.c_float]),
    "gd_near_r2": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float]),
}

# Helpers exported beside the header API (not part of include/geauxdock_amd.h).
EXTRA_PROTOTYPES = {
    "gd_host_job_diag": (C.c_int, [_vp, _ip, _fp, C.c_int, _ip, C.c_int]),
}


def declare(lib, names=None, optional=False):
    """Attach prototypes to a ctypes CDLL; raises AttributeError if missing."""
    table = dict(PROTOTYPES)
    table.update(EXTRA_PROTOTYPES)
    for name, (res, args) in table.items():
        if names is not None and name not in names:
            continue
        try:
            fn = getattr(lib, name)
        except AttributeError:
            if optional:
                continue
            raise
        fn.restype = res
        fn.argtypes = args
    return lib
