"""ctypes mirrors of the boundary structs (include/jurassic_types.h).

The layouts follow the reference's src/jurassic.h:228-570 member for member;
``check_layout`` compares the sizes with the ones the C header pins.
"""
import ctypes as C

LEN = 5000
NWMAX = 5
NDMAX = 130
NGMAX = 25
NLOS = 10000
NPMAX = 1000
NRMAX = 1000
SCAMOD = 30
NLMAX = 10
NTHETA = 181


class queue_t(C.Structure):
    _fields_ = [("items", C.c_void_p), ("capacity", C.c_int), ("begin", C.c_int),
                ("end", C.c_int), ("state", C.c_int)]


class ctl_t(C.Structure):
    _fields_ = [
        ("ng", C.c_int), ("emitter", (C.c_char * LEN) * NGMAX), ("nd", C.c_int),
        ("nw", C.c_int), ("nu", C.c_double * NDMAX), ("window", C.c_int * NDMAX),
        ("tblbase", C.c_char * LEN), ("hydz", C.c_double), ("ctm_co2", C.c_int),
        ("ctm_h2o", C.c_int), ("ctm_n2", C.c_int), ("ctm_o2", C.c_int),
        ("sca_n", C.c_int), ("sca_mult", C.c_int), ("sca_ext", C.c_char * LEN),
        ("ip", C.c_int), ("cz", C.c_double), ("cx", C.c_double), ("refrac", C.c_int),
        ("rayds", C.c_double), ("raydz", C.c_double), ("transs", C.c_double),
        ("fov", C.c_char * LEN), ("retp_zmin", C.c_double), ("retp_zmax", C.c_double),
        ("rett_zmin", C.c_double), ("rett_zmax", C.c_double),
        ("retq_zmin", C.c_double * NGMAX), ("retq_zmax", C.c_double * NGMAX),
        ("retk_zmin", C.c_double * NWMAX), ("retk_zmax", C.c_double * NWMAX),
        ("retnn_zmin", C.c_double), ("retnn_zmax", C.c_double), ("retnn", C.c_int),
        ("retrr", C.c_int), ("retss", C.c_int), ("write_bbt", C.c_int),
        ("write_matrix", C.c_int), ("queue", queue_t),
    ]


class los_t(C.Structure):
    _fields_ = [
        ("np", C.c_int), ("z", C.c_double * NLOS), ("lon", C.c_double * NLOS),
        ("lat", C.c_double * NLOS), ("p", C.c_double * NLOS), ("t", C.c_double * NLOS),
        ("q", (C.c_double * NGMAX) * NLOS), ("k", (C.c_double * NWMAX) * NLOS),
        ("aeroi", C.c_int * NLOS), ("aerofac", C.c_double * NLOS), ("tsurf", C.c_double),
        ("ds", C.c_double * NLOS), ("u", (C.c_double * NGMAX) * NLOS),
    ]


class obs_t(C.Structure):
    _fields_ = [
        ("nr", C.c_int), ("time", C.c_double * NRMAX), ("obsz", C.c_double * NRMAX),
        ("obslon", C.c_double * NRMAX), ("obslat", C.c_double * NRMAX),
        ("vpz", C.c_double * NRMAX), ("vplon", C.c_double * NRMAX),
        ("vplat", C.c_double * NRMAX), ("tpz", C.c_double * NRMAX),
        ("tplon", C.c_double * NRMAX), ("tplat", C.c_double * NRMAX),
        ("tau", (C.c_double * NRMAX) * NDMAX), ("rad", (C.c_double * NRMAX) * NDMAX),
    ]


class aero_t(C.Structure):
    _fields_ = [
        ("nm", C.c_int), ("top_mod", C.c_double * SCAMOD), ("bottom_mod", C.c_double * SCAMOD),
        ("trans_mod", C.c_double * SCAMOD), ("type", (C.c_char * LEN) * SCAMOD),
        ("filepath", (C.c_char * LEN) * SCAMOD), ("nn", C.c_double * SCAMOD),
        ("rr", C.c_double * SCAMOD), ("ss", C.c_double * SCAMOD), ("nl", C.c_int),
        ("nmod", C.c_int * NLMAX), ("top", C.c_double * NLMAX), ("bottom", C.c_double * NLMAX),
        ("trans", C.c_double * NLMAX), ("beta_e", (C.c_double * NDMAX) * NLMAX),
        ("beta_s", (C.c_double * NDMAX) * NLMAX), ("beta_a", (C.c_double * NDMAX) * NLMAX),
        ("p", ((C.c_double * NTHETA) * NDMAX) * NLMAX),
    ]


class atm_t(C.Structure):
    _fields_ = [
        ("np", C.c_int), ("time", C.c_double * NPMAX), ("z", C.c_double * NPMAX),
        ("lon", C.c_double * NPMAX), ("lat", C.c_double * NPMAX), ("p", C.c_double * NPMAX),
        ("t", C.c_double * NPMAX), ("q", (C.c_double * NPMAX) * NGMAX),
        ("k", (C.c_double * NPMAX) * NWMAX), ("init", C.c_int),
    ]


class ret_t(C.Structure):
    _fields_ = [
        ("dir", C.c_char * LEN), ("kernel_recomp", C.c_int), ("conv_itmax", C.c_int),
        ("conv_dmin", C.c_double), ("resmax", C.c_double), ("err_ana", C.c_int),
        ("err_formod", C.c_double * NDMAX), ("err_noise", C.c_double * NDMAX),
        ("err_press", C.c_double), ("err_press_cz", C.c_double), ("err_press_ch", C.c_double),
        ("err_temp", C.c_double), ("err_temp_cz", C.c_double), ("err_temp_ch", C.c_double),
        ("err_nn", C.c_double), ("err_rr", C.c_double), ("err_ss", C.c_double),
        ("err_q", C.c_double * NGMAX), ("err_q_cz", C.c_double * NGMAX), ("err_q_ch", C.c_double * NGMAX),
        ("err_k", C.c_double * NWMAX), ("err_k_cz", C.c_double * NWMAX), ("err_k_ch", C.c_double * NWMAX),
    ]


SIZES = {"queue_t": 24, "ctl_t": 142240, "atm_t": 288016, "obs_t": 2160008,
         "los_t": 5000016, "aero_t": 2215336, "ret_t": 7904}


def check_layout():
    for name, size in SIZES.items():
        got = C.sizeof(globals()[name])
        if got != size:
            raise AssertionError("%s: sizeof %d != %d" % (name, got, size))
    assert ctl_t.queue.offset == 142216
    assert obs_t.rad.offset == 1120008
    assert aero_t.p.offset == 332936


def argv(ctlfile, *extra):
    """Build a C argv array like the reference CLI's: prog, ctl, then KEY VALUE pairs."""
    items = [b"jurassic", str(ctlfile).encode()] + [str(x).encode() for x in extra]
    arr = (C.c_char_p * (len(items) + 1))(*items, None)
    return len(items), arr
