"""ctypes mirror of include/emd_hip.h and include/emd_app.h."""
import ctypes as C

EMD_OK = 0
EMD_FP64, EMD_MIXED = 0, 1
EMD_NEIGH_CSR, EMD_NEIGH_2D = 0, 1

P = C.POINTER
vp = C.c_void_p
dptr = P(C.c_double)
iptr = P(C.c_int)


class EmdSystem(C.Structure):
    _fields_ = [("x", dptr), ("v", dptr), ("f", dptr), ("f32", P(C.c_float)),
                ("type", iptr), ("id", iptr), ("q", dptr), ("mass", dptr),
                ("n_local", C.c_int), ("n_ghost", C.c_int), ("n_max", C.c_int),
                ("ntypes", C.c_int), ("n_global", C.c_int),
                ("domain", C.c_double * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("boltz", C.c_double), ("mvv2e", C.c_double), ("dt", C.c_double)]


class EmdVec3View(C.Structure):
    _fields_ = [("data", dptr), ("layout", C.c_int), ("ld", C.c_long)]


EMD_LAYOUT_RIGHT, EMD_LAYOUT_LEFT = 0, 1


class EmdHalo(C.Structure):
    _fields_ = [("num_ghost", C.c_int * 6), ("pack", iptr * 6), ("cap", C.c_int * 6),
                ("root", iptr), ("image", iptr), ("root_cap", C.c_int), ("n_ghost", C.c_int)]


class EmdBins(C.Structure):
    _fields_ = [("nbinx", C.c_int), ("nbiny", C.c_int), ("nbinz", C.c_int), ("nhalo", C.c_int),
                ("minx", C.c_double), ("maxx", C.c_double), ("miny", C.c_double),
                ("maxy", C.c_double), ("minz", C.c_double), ("maxz", C.c_double),
                ("nbins", C.c_int), ("n_range", C.c_int),
                ("bincount", iptr), ("binoffsets", iptr), ("permute", iptr), ("binid", iptr),
                ("cap_bins", C.c_int), ("cap_range", C.c_int), ("max_count", C.c_int)]


class EmdNeigh(C.Structure):
    _fields_ = [("layout", C.c_int), ("n", C.c_int), ("total", C.c_long),
                ("maxneighs", C.c_int), ("row", iptr), ("num", iptr), ("entries", iptr),
                ("cap_entries", C.c_long), ("cap_rows", C.c_int)]


class EmdLJParams(C.Structure):
    _fields_ = [("ntypes", C.c_int), ("lj1", (C.c_double * 16) * 16),
                ("lj2", (C.c_double * 16) * 16), ("cutsq", (C.c_double * 16) * 16),
                ("table", C.POINTER(C.c_double))]


class EmdLJIDialParams(C.Structure):
    _fields_ = [("lj", EmdLJParams), ("intensity", (C.c_double * 16) * 16)]


class EmdSnapParams(C.Structure):
    _fields_ = [("twojmax", C.c_int), ("rfac0", C.c_double), ("rmin0", C.c_double),
                ("rcutfac", C.c_double), ("switchflag", C.c_int), ("bzeroflag", C.c_int),
                ("radelem", C.c_double), ("wjelem", C.c_double), ("ncoeff", C.c_int),
                ("coeff", dptr)]


class EmdBatchDesc(C.Structure):
    _fields_ = [("nrep", C.c_int), ("natoms", iptr), ("box", dptr), ("mass", dptr),
                ("eps", C.c_double), ("sigma", C.c_double), ("cut", C.c_double),
                ("skin", C.c_double), ("dt", C.c_double), ("rebuild_every", C.c_int),
                ("precision", C.c_int)]


HIP_PROTOS = {
    "emd_create": (C.c_int, [C.c_int, P(vp)]),
    "emd_destroy": (C.c_int, [vp]),
    "emd_last_error": (C.c_char_p, [vp]),
    "emd_synchronize": (C.c_int, [vp]),
    "emd_get_stream": (vp, [vp]),
    "emd_version": (C.c_char_p, []),
    "emd_malloc": (C.c_int, [vp, P(vp), C.c_size_t]),
    "emd_free": (C.c_int, [vp, vp]),
    "emd_memcpy_h2d": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "emd_memcpy_d2h": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "emd_memcpy_d2d": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "emd_profile_enable": (C.c_int, [vp, C.c_int]),
    "emd_profile_reset": (C.c_int, [vp]),
    "emd_profile_select": (C.c_int, [vp, C.c_char_p]),
    "emd_profile_get": (C.c_int, [vp, C.c_char_p, P(C.c_double), P(C.c_long)]),
    "emd_profile_report": (C.c_size_t, [vp, C.c_char_p, C.c_size_t]),
    "emd_system_grow": (C.c_int, [vp, P(EmdSystem), C.c_int]),
    "emd_system_free": (C.c_int, [vp, P(EmdSystem)]),
    "emd_system_upload": (C.c_int, [vp, P(EmdSystem), C.c_int, dptr, dptr, iptr, iptr, dptr]),
    "emd_system_download": (C.c_int, [vp, P(EmdSystem), C.c_int, dptr, dptr, dptr, iptr, iptr,
                                      dptr]),
    "emd_system_set_mass": (C.c_int, [vp, P(EmdSystem), C.c_int, dptr]),
    "emd_system_import": (C.c_int, [vp, P(EmdSystem), C.c_int, EmdVec3View, EmdVec3View,
                                    EmdVec3View, vp, vp, vp]),
    "emd_system_export": (C.c_int, [vp, P(EmdSystem), C.c_int, EmdVec3View, EmdVec3View,
                                    EmdVec3View]),
    "emd_ctx_shared": (C.c_int, [C.c_int, P(vp)]),
    "emd_bins_import": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_double),
                                  P(C.c_double), vp, vp, C.c_int, vp, C.c_int, P(EmdBins)]),
    "emd_comm_exchange": (C.c_int, [vp, P(EmdSystem)]),
    "emd_comm_exchange_halo": (C.c_int, [vp, P(EmdSystem), C.c_double, P(EmdHalo)]),
    "emd_comm_update_halo": (C.c_int, [vp, P(EmdSystem), P(EmdHalo)]),
    "emd_comm_update_force": (C.c_int, [vp, P(EmdSystem), P(EmdHalo), C.c_int]),
    "emd_halo_free": (C.c_int, [vp, P(EmdHalo)]),
    "emd_binning": (C.c_int, [vp, P(EmdSystem), C.c_double, C.c_double, C.c_double, C.c_int,
                              C.c_int, C.c_int, C.c_int, P(EmdBins)]),
    "emd_bins_free": (C.c_int, [vp, P(EmdBins)]),
    "emd_neigh_build": (C.c_int, [vp, P(EmdSystem), P(EmdBins), C.c_double, C.c_int, C.c_int,
                                  C.c_int, P(EmdNeigh)]),
    "emd_neigh_free": (C.c_int, [vp, P(EmdNeigh)]),
    "emd_force_zero": (C.c_int, [vp, P(EmdSystem), C.c_int]),
    "emd_lj_force": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), P(EmdLJParams), C.c_int, C.c_int]),
    "emd_lj_force_store": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), P(EmdLJParams), C.c_int]),
    "emd_lj_idial_force": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), P(EmdLJIDialParams),
                                     C.c_int, C.c_int]),
    "emd_lj_idial_force_store": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh),
                                           P(EmdLJIDialParams), C.c_int]),
    "emd_lj_energy": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), P(EmdLJParams), C.c_int,
                                P(C.c_double)]),
    "emd_lj_energy_atoms": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), P(EmdLJParams), C.c_int,
                                      dptr]),
    "emd_lj_energy_noalloc": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), P(EmdLJParams),
                                        C.c_int, P(C.c_double)]),
    "emd_lj_set_lanes": (C.c_int, [vp, C.c_int, C.c_int]),
    "emd_nve_initial": (C.c_int, [vp, P(EmdSystem), C.c_double, C.c_double, C.c_int]),
    "emd_nve_final": (C.c_int, [vp, P(EmdSystem), C.c_double, C.c_int]),
    "emd_nve_final_initial": (C.c_int, [vp, P(EmdSystem), C.c_double, C.c_double, C.c_int]),
    "emd_nve_step_boundary": (C.c_int, [vp, P(EmdSystem), P(EmdHalo), C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_int]),
    "emd_reduce_mv2": (C.c_int, [vp, P(EmdSystem), P(C.c_double)]),
    "emd_snap_create": (C.c_int, [vp, P(EmdSnapParams), P(vp)]),
    "emd_snap_destroy": (C.c_int, [vp, vp]),
    "emd_snap_set_coeff": (C.c_int, [vp, vp, dptr]),
    "emd_snap_static_geometry": (C.c_int, [vp, vp, C.c_int]),
    "emd_snap_force": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), vp, C.c_int]),
    "emd_snap_force_zeroed": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), vp, C.c_int]),
    "emd_snap_energy": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), vp, P(C.c_double), dptr]),
    "emd_snap_adjoint_beta": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), P(EmdHalo), vp, dptr,
                                        C.c_double, C.c_int, dptr]),
    "emd_snap_descriptors": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), vp, dptr]),
    "emd_lj_adjoint": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), C.c_double, C.c_double,
                                 C.c_double, dptr, C.c_double, C.c_int, dptr]),
    "emd_lj_adjoint_atoms": (C.c_int, [vp, P(EmdSystem), P(EmdNeigh), C.c_double, C.c_double,
                                       C.c_double, dptr, dptr, C.c_int, dptr]),
    "emd_batch_create": (C.c_int, [vp, P(EmdBatchDesc), P(vp)]),
    "emd_batch_destroy": (C.c_int, [vp, vp]),
    "emd_batch_upload": (C.c_int, [vp, vp, dptr, dptr]),
    "emd_batch_download": (C.c_int, [vp, vp, dptr, dptr, dptr]),
    "emd_batch_setup": (C.c_int, [vp, vp]),
    "emd_batch_run": (C.c_int, [vp, vp, C.c_int, C.c_int]),
    "emd_batch_thermo": (C.c_int, [vp, vp, dptr, dptr]),
    "emd_batch_set_tile": (C.c_int, [vp, vp, C.c_int, C.c_int]),
    "emd_batch_autotune": (C.c_int, [vp, vp, C.c_int, iptr, iptr, P(C.c_double)]),
    "emd_batch_info": (C.c_int, [vp, vp, P(C.c_long), iptr, iptr]),
    "emd_batch_neigh": (C.c_int, [vp, vp, iptr, P(C.c_uint)]),
}

APP_PROTOS = {
    "emd_app_create": (C.c_int, [C.c_int, P(C.c_char_p), C.c_int, P(vp)]),
    "emd_app_init": (C.c_int, [vp]),
    "emd_app_run": (C.c_int, [vp, C.c_int]),
    "emd_app_set_phase_sync": (C.c_int, [vp, C.c_int]),
    "emd_app_set_fused_integrate": (C.c_int, [vp, C.c_int]),
    "emd_app_destroy": (C.c_int, [vp]),
    "emd_app_nsteps": (C.c_int, [vp]),
    "emd_app_natoms": (C.c_int, [vp]),
    "emd_app_nlocal": (C.c_int, [vp]),
    "emd_app_nghost": (C.c_int, [vp]),
    "emd_app_thermo": (C.c_int, [vp, P(C.c_double)]),
    "emd_app_time": (C.c_double, [vp, C.c_int]),
    "emd_app_download": (C.c_int, [vp, dptr, dptr, dptr, iptr, iptr, dptr]),
    "emd_app_dump": (C.c_int, [vp, C.c_char_p]),
    "emd_app_dump_neigh": (C.c_int, [vp, C.c_char_p]),
    "emd_app_ctx": (vp, [vp]),
    "emd_app_system": (P(EmdSystem), [vp]),
    "emd_app_neigh": (P(EmdNeigh), [vp]),
    "emd_app_halo": (P(EmdHalo), [vp]),
    "emd_app_precision": (C.c_int, [vp]),
    "emd_replica_loop_run": (C.c_int, [vp, P(EmdBatchDesc), dptr, dptr, C.c_int, P(C.c_double),
                                       dptr, dptr]),
}


def _attach(lib, protos):
    for name, (res, args) in protos.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args


def attach_hip(lib):
    _attach(lib, HIP_PROTOS)


def attach_app(lib):
    _attach(lib, APP_PROTOS)
