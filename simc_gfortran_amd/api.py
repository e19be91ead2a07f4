"""Host-side interface to libsimc_gpu.so (the MI355X spectrometer path).

Mirrors the reference's per-arm subroutines, vectorised over events held in
GPU memory (torch float64 tensors on the current HIP device):

    mc_hms(p_spec, th_spec, dpp, x, y, z, dxdz, dydz, ..., m2, ms_flag,
           wcs_flag, decay_flag, fry, pathlen, using_HMScoll)   hms/mc_hms.f:1-4
    mc_shms(..., spectr, using_SHMScoll)                        shms/mc_shms.f:1-4
    mc_hms_recon / mc_shms_recon                                hms/mc_hms_recon.f:1

The compute runs only in the HIP library; importing this module on a machine
without the built library raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libsimc_gpu.so")
MAP_DIR = os.path.join(REPO, "data", "maps")

HMS = 1
SOS = 2
HRSR = 3
HRSL = 4
SHMS = 5
RNG_TAPE = 0
RNG_PHILOX = 1
FIT_REF_F32 = 1
FIT_FP64 = 0

HMS_COUNTERS = [
    "trials", "slit_hor", "slit_vert", "slit_oct", "Q1_in", "Q1_mid", "Q1_out", "Q2_in",
    "Q2_mid", "Q2_out", "Q3_in", "Q3_mid", "Q3_out", "D1_in", "D1_out", "hut", "dc1", "dc2",
    "scin", "cal", "successes", "coll"]
SHMS_COUNTERS = [
    "trials", "successes", "targ_hor", "targ_vert", "targ_oct", "slit_hor", "slit_vert",
    "slit_oct", "HB_in", "HB_men", "HB_mex", "HB_out", "Q1_in", "Q1_men", "Q1_mid", "Q1_mex",
    "Q1_out", "Q2_in", "Q2_men", "Q2_mid", "Q2_mex", "Q2_out", "Q3_in", "Q3_men", "Q3_mid",
    "Q3_mex", "Q3_out", "D1_in", "D1_flr", "D1_men", "D1_mid1", "D1_mid2", "D1_mid3",
    "D1_mid4", "D1_mid5", "D1_mid6", "D1_mid7", "D1_mex", "D1_out", "BP_in", "BP_out", "hut",
    "dc1", "dc2", "cer1", "cer2", "s1", "s2", "s3", "cal", "stop_id", "coll"]
# common /spectrometer2/ (sos/struct_sos.inc:9-17); "exit" sums the three exit-aperture
# stops (sos/mc_sos.f:343-358), "scin" counts events lost at the trigger.
SOS_COUNTERS = [
    "trials", "slit_hor", "slit_vert", "slit_oct", "quad_in", "quad_mid", "quad_out", "bm01_in",
    "bm01_out", "bm02_in", "bm02_out", "exit", "hut", "dc1", "dc2", "scin", "cal", "successes"]
# common /spectrometer3/ (HRS-R) and /spectrometer4/ (HRS-L), hrsr/struct_hrsr.inc:24-40
HRS_COUNTERS = [
    "trials", "slit_hor", "slit_vert", "Q1_in", "Q1_mid", "Q1_out", "Q2_in", "Q2_mid", "Q2_out",
    "Q3_in", "Q3_mid", "Q3_out", "D1_in", "D1_out", "hut", "dc1", "dc2", "s1", "s2", "cal",
    "successes"]
COUNTERS = {HMS: HMS_COUNTERS, SOS: SOS_COUNTERS, HRSR: HRS_COUNTERS, HRSL: HRS_COUNTERS,
            SHMS: SHMS_COUNTERS}

# Every exported C entry point (include/simc_gpu.h).
EXPORTS = [
    "simc_ctx_create", "simc_ctx_create_files", "simc_ctx_load_arm", "simc_ctx_destroy",
    "simc_hms_run", "simc_shms_run", "simc_sos_run", "simc_hrsr_run", "simc_hrsl_run", "simc_arm_run",
    "simc_arm_run_noalloc", "simc_arm_run_timed", "simc_recon_run", "simc_counters_get",
    "simc_counters_reset", "simc_map_class_info", "simc_poly_flops_full", "simc_stop_flops",
    "simc_stop_flops_stage", "simc_stop_flops_hut", "simc_arm_run_host", "simc_recon_run_host",
    "simc_batch_create", "simc_batch_run", "simc_batch_counters", "simc_batch_set_chunk",
    "simc_batch_tune", "simc_batch_destroy", "simc_batch_passes", "simc_adjoint_run", "simc_recon_coefficients",
    "simc_target_prep", "simc_target_prep_host", "simc_gen_hep", "simc_gen_hep_host", "simc_gen_hep_rad",
    "simc_field_create", "simc_field_destroy", "simc_track_from_tgt", "simc_track_to_tgt",
    "simc_track_from_tgt_host", "simc_track_to_tgt_host",
    "simc_ntuple_open", "simc_ntuple_write", "simc_ntuple_write_device", "simc_ntuple_close", "simc_stop_report",
    "simc_last_error",
    "simc_version",
]

_vp = C.c_void_p


class Settings(C.Structure):
    _fields_ = [("p_spec", C.c_double), ("m2", C.c_double), ("ctau", C.c_double),
                ("ms_flag", C.c_int32), ("wcs_flag", C.c_int32), ("decay_flag", C.c_int32),
                ("using_coll", C.c_int32), ("fit_mode", C.c_int32), ("sieve", C.c_int32),
                ("wire_sigma", C.c_double)]


class EventsIn(C.Structure):
    _fields_ = [(n, _vp) for n in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")]


OUT_FIELDS = ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen",
              "resmult", "m2", "ok", "stop", "used")
# the outputs a run always writes (resmult, m2 and used are optional)
RUN_REQUIRED = ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "ok", "stop")


class EventsOut(C.Structure):
    _fields_ = [(n, _vp) for n in OUT_FIELDS]


class Rng(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("seed", C.c_uint64),
                ("event0", C.c_uint64), ("tape", _vp), ("stride", C.c_int64)]


class Setting(C.Structure):
    _fields_ = [("arm", C.c_int32), ("reserved", C.c_int32), ("s", Settings), ("seed", C.c_uint64),
                ("event0", C.c_uint64), ("offset", C.c_int64), ("count", C.c_int64)]


class Target(C.Structure):
    """simc_target: the target of trip_thru_target (X0 <= 0: init.f's value)."""
    _fields_ = [("A", C.c_double), ("Z", C.c_double), ("rho", C.c_double), ("X0", C.c_double),
                ("length", C.c_double), ("angle", C.c_double), ("zoffset", C.c_double),
                ("can", C.c_int32), ("reserved", C.c_int32)]


class PrepSettings(C.Structure):
    """simc_prep_settings: one particle's preparation (narm 1 beam, 2 electron, 3 hadron)."""
    _fields_ = [("narm", C.c_int32), ("spectrometer", C.c_int32), ("using_eloss", C.c_int32),
                ("mc_smear", C.c_int32), ("p_spec", C.c_double), ("th_spec", C.c_double),
                ("phi", C.c_double), ("mass", C.c_double), ("coulomb", C.c_double),
                ("beam_energy", C.c_double), ("offset", C.c_double * 5)]


VERTEX_FIELDS = ("x", "y", "z", "E", "p", "E_vertex", "theta", "delta", "xptar", "yptar", "dang_yp",
                 "dang_xp")
PREP_OUT_FIELDS = ("dpp", "x", "y", "z", "dxdz", "dydz", "fry", "eloss", "teff", "dang1", "dang2", "used")


class VertexIn(C.Structure):
    _fields_ = [(n, _vp) for n in VERTEX_FIELDS]


class PrepOut(C.Structure):
    _fields_ = [(n, _vp) for n in PREP_OUT_FIELDS]


class HepSettings(C.Structure):
    """simc_hep_settings: H(e,e'p) elastic generation (event.f generate,
    complete_ev, complete_main for doing_hyd_elast, radiation off)."""
    _fields_ = [("Ebeam", C.c_double), ("dEbeam", C.c_double), ("xwid", C.c_double), ("ywid", C.c_double),
                ("xoffset", C.c_double), ("yoffset", C.c_double), ("fr1", C.c_double), ("fr2", C.c_double),
                ("coulomb", C.c_double), ("e_yptar", C.c_double * 2), ("e_xptar", C.c_double * 2),
                ("e_theta", C.c_double), ("e_phi", C.c_double), ("e_P", C.c_double),
                ("p_theta", C.c_double), ("p_phi", C.c_double), ("p_P", C.c_double), ("Mh", C.c_double),
                ("fr_pattern", C.c_int32), ("using_eloss", C.c_int32), ("using_coulomb", C.c_int32),
                ("reserved", C.c_int32)]


def hep_settings(**kw):
    """HepSettings from keywords (e_yptar, e_xptar as (min, max))."""
    h = HepSettings()
    for k, v in kw.items():
        if k in ("e_yptar", "e_xptar"):
            getattr(h, k)[0], getattr(h, k)[1] = v
        elif k in ("fr_pattern", "using_eloss", "using_coulomb"):
            setattr(h, k, int(v))
        else:
            setattr(h, k, float(v))
    return h


HEP_OUT_FIELDS = ("x", "y", "z", "rasterx", "rastery", "Ein", "eloss", "teff",
                  "e_xptar", "e_yptar", "e_theta", "e_phi", "e_E", "e_delta",
                  "p_xptar", "p_yptar", "p_theta", "p_phi", "p_P", "p_E", "p_delta",
                  "Q2", "nu", "q", "jacobian", "sigcc", "weight")


class HepOut(C.Structure):
    _fields_ = [(n, _vp) for n in HEP_OUT_FIELDS] + [("ok", _vp), ("used", _vp)]


class RadSettings(C.Structure):
    """simc_rad_settings: the radiative options and run constants that
    generate_rad reads (radc.f; radc_init and limits_init of init.f)."""
    _fields_ = [("rad_flag", C.c_int32), ("extrad_flag", C.c_int32), ("intcor_mode", C.c_int32),
                ("use_expon", C.c_int32), ("use_offshell_rad", C.c_int32), ("doing_tail", C.c_int32 * 3),
                ("e_arm", C.c_int32), ("p_arm", C.c_int32), ("dE_edge_test", C.c_double),
                ("Egamma_gen_max", C.c_double), ("edge_e_E", C.c_double * 2), ("edge_p_E", C.c_double * 2),
                ("edge_Em", C.c_double * 2), ("vertex_Trec", C.c_double * 2), ("Egamma_tot_max", C.c_double),
                ("etatzai", C.c_double)]


def rad_settings(**kw):
    """RadSettings from keywords (pairs as (min, max), doing_tail as three
    flags); etatzai absent or <= 0: from the target's Z."""
    r = RadSettings()
    for k, v in kw.items():
        if k == "doing_tail":
            for i in range(3):
                r.doing_tail[i] = int(v[i])
        elif k in ("edge_e_E", "edge_p_E", "edge_Em", "vertex_Trec"):
            getattr(r, k)[0], getattr(r, k)[1] = v
        elif k in ("rad_flag", "extrad_flag", "intcor_mode", "use_expon", "use_offshell_rad", "e_arm", "p_arm"):
            setattr(r, k, int(v))
        elif k in ("dE_edge_test", "Egamma_gen_max", "Egamma_tot_max", "etatzai"):
            setattr(r, k, float(v))
        elif k != "targ_A":  # the target carries A
            raise KeyError(k)
    return r


RAD_OUT_FIELDS = ("Egamma1", "Egamma2", "Egamma3", "gen_weight", "orig_Ein", "orig_e_E", "orig_e_delta",
                  "orig_p_E", "orig_p_P", "orig_p_delta", "eloss2", "teff2", "eloss3", "teff3", "hardcorfac")


class RadOut(C.Structure):
    _fields_ = [(n, _vp) for n in RAD_OUT_FIELDS] + [("ntail", _vp)]


class SimcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"simc error {code}: {msg}")
        self.code = code


_lib = None


def library(path: str = LIB_PATH):
    """Load libsimc_gpu.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m simc_gfortran_amd.build`")
    L = C.CDLL(path)
    L.simc_ctx_create.argtypes = [C.c_char_p, C.POINTER(_vp)]
    L.simc_ctx_create_files.argtypes = [C.c_char_p] * 4 + [C.POINTER(_vp)]
    L.simc_ctx_load_arm.argtypes = [_vp, C.c_int, C.c_char_p, C.c_char_p]
    L.simc_ctx_destroy.argtypes = [_vp]
    L.simc_ctx_destroy.restype = None
    L.simc_arm_run.argtypes = [_vp, C.c_int, C.POINTER(Settings), C.POINTER(EventsIn),
                               C.POINTER(EventsOut), C.c_int64, C.POINTER(Rng), _vp]
    L.simc_arm_run_noalloc.argtypes = [_vp, C.c_int, C.POINTER(Settings), C.POINTER(EventsIn),
                                       C.POINTER(EventsOut), C.c_int64, C.POINTER(Rng), _vp]
    L.simc_arm_run_timed.argtypes = [_vp, C.c_int, C.POINTER(Settings), C.POINTER(EventsIn),
                                     C.POINTER(EventsOut), C.c_int64, C.POINTER(Rng), _vp,
                                     C.POINTER(_vp)]
    L.simc_recon_run.argtypes = [_vp, C.c_int, C.c_int64] + [_vp] * 9 + [_vp]
    L.simc_batch_create.argtypes = [_vp, C.POINTER(Setting), C.c_int, C.POINTER(_vp)]
    L.simc_batch_run.argtypes = [_vp, C.POINTER(EventsIn), C.POINTER(EventsOut), _vp]
    L.simc_batch_counters.argtypes = [_vp, C.c_int, C.POINTER(C.c_int64), C.c_int]
    L.simc_batch_set_chunk.argtypes = [_vp, C.c_int64]
    L.simc_batch_tune.argtypes = [_vp, C.POINTER(EventsIn), C.POINTER(EventsOut), _vp,
                                  C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.simc_adjoint_run.argtypes = [_vp, C.c_int, C.POINTER(Settings), C.POINTER(EventsIn), _vp, C.c_int64,
                                   C.POINTER(Rng), _vp, C.POINTER(C.c_double), _vp, _vp, _vp]
    L.simc_recon_coefficients.argtypes = [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.simc_gen_hep.argtypes = [_vp, C.POINTER(Target), C.POINTER(HepSettings), C.POINTER(HepOut), C.c_int64,
                               C.POINTER(Rng), _vp]
    L.simc_gen_hep_rad.argtypes = [_vp, C.POINTER(Target), C.POINTER(HepSettings), C.POINTER(RadSettings),
                                   C.POINTER(HepOut), C.POINTER(RadOut), C.c_int64, C.POINTER(Rng), _vp]
    L.simc_gen_hep_host.argtypes = [_vp, C.POINTER(Target), C.POINTER(HepSettings), C.POINTER(HepOut), C.c_int64,
                                    C.POINTER(Rng)]
    L.simc_target_prep.argtypes = [_vp, C.POINTER(Target), C.POINTER(PrepSettings), C.POINTER(VertexIn),
                                   C.POINTER(PrepOut), C.c_int64, C.POINTER(Rng), _vp]
    L.simc_target_prep_host.argtypes = [_vp, C.POINTER(Target), C.POINTER(PrepSettings), C.POINTER(VertexIn),
                                        C.POINTER(PrepOut), C.c_int64, C.POINTER(Rng)]
    L.simc_field_create.argtypes = [C.c_char_p] + [C.c_double] * 4 + [C.POINTER(_vp)]
    L.simc_field_destroy.argtypes = [_vp]
    L.simc_field_destroy.restype = None
    L.simc_track_from_tgt.argtypes = [_vp, C.c_int, C.c_int64] + [_vp] * 6 + [C.c_double, _vp, _vp]
    L.simc_track_to_tgt.argtypes = ([_vp, _vp, C.c_int, C.c_int, C.c_int64] + [_vp] * 11 +
                                    [C.c_double] * 3 + [_vp, _vp])
    L.simc_track_from_tgt_host.argtypes = [_vp, C.c_int, C.c_int64] + [_vp] * 6 + [C.c_double, _vp]
    L.simc_track_to_tgt_host.argtypes = ([_vp, _vp, C.c_int, C.c_int, C.c_int64] + [_vp] * 11 +
                                         [C.c_double] * 3 + [_vp])
    L.simc_ntuple_open.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(_vp)]
    L.simc_ntuple_write.argtypes = [_vp, _vp, C.c_int64]
    L.simc_ntuple_write_device.argtypes = [_vp, C.POINTER(_vp), _vp, C.c_int64, C.POINTER(C.c_int64), _vp]
    L.simc_ntuple_close.argtypes = [_vp]
    L.simc_stop_report.argtypes = [_vp, C.c_char_p, C.c_int, C.c_int]
    L.simc_batch_destroy.argtypes = [_vp]
    L.simc_batch_destroy.restype = None
    L.simc_batch_passes.argtypes = [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.simc_arm_run_host.argtypes = [_vp, C.c_int, C.POINTER(Settings), C.POINTER(EventsIn),
                                    C.POINTER(EventsOut), C.c_int64, C.POINTER(Rng)]
    L.simc_recon_run_host.argtypes = [_vp, C.c_int, C.c_int64] + [_vp] * 9
    L.simc_counters_get.argtypes = [_vp, C.c_int, C.POINTER(C.c_int64), C.c_int]
    L.simc_counters_reset.argtypes = [_vp, C.c_int]
    L.simc_map_class_info.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.simc_poly_flops_full.argtypes = [_vp, C.c_int]
    L.simc_poly_flops_full.restype = C.c_int64
    L.simc_stop_flops.argtypes = [_vp, C.c_int, C.POINTER(C.c_int64), C.c_int]
    L.simc_stop_flops_stage.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int]
    L.simc_stop_flops_hut.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int]
    L.simc_last_error.restype = C.c_char_p
    L.simc_version.restype = C.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise SimcError(rc, library().simc_last_error().decode())


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _dev_f64(t, n, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise TypeError(f"{name}: expected a float64 tensor on a HIP device")
    if t.numel() != n or not t.is_contiguous():
        raise ValueError(f"{name}: expected {n} contiguous elements")
    return t


# dtype of every event output column (simc_events_out)
OUT_DTYPES = {k: torch.float64 for k in ("dpp", "y", "dxdz", "dydz", "x_fp", "dx_fp", "y_fp", "dy_fp",
                                          "pathlen", "resmult", "m2")}
OUT_DTYPES.update(ok=torch.uint8, stop=torch.uint8, used=torch.int32)


def _dev_tensor(t, n, name, dtype, at_least=False):
    """A device tensor of `dtype` with n contiguous elements (or at least n):
    what a kernel may address through its pointer."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dtype:
        raise TypeError(f"{name}: expected a {dtype} tensor on a HIP device")
    if not t.is_contiguous() or (t.numel() < n if at_least else t.numel() != n):
        raise ValueError(f"{name}: expected {'at least ' if at_least else ''}{n} contiguous elements")
    return t


def _check_outputs(out, n, required, at_least=False):
    """Validate caller-supplied event outputs: every present column has its
    dtype, lives on a HIP device and holds n elements; `required` ones must
    be present."""
    for k in required:
        if out.get(k) is None:
            raise ValueError(f"out: column {k} is required")
    for k, dt in OUT_DTYPES.items():
        if out.get(k) is not None:
            _dev_tensor(out[k], n, "out[" + k + "]", dt, at_least)


def _check_tape(tape, n, stride):
    if stride <= 0:
        raise ValueError("tape rng needs a positive stride")
    _dev_tensor(tape, n * int(stride), "tape", torch.float64, at_least=True)


class Spectrometers:
    """A context: the COSY maps of the arms on the device plus stop counters
    (transp_init and the recon readers, shared/transp.f:294-478).  HMS and SHMS
    maps are required; SOS, HRS-R and HRS-L maps are loaded when map_dir has
    them."""

    def __init__(self, map_dir: str = MAP_DIR, files=None):
        L = library()
        h = _vp()
        if files is not None:
            _check(L.simc_ctx_create_files(*[f.encode() for f in files], C.byref(h)))
        else:
            _check(L.simc_ctx_create(map_dir.encode(), C.byref(h)))
        self._h = h

    def load_arm(self, arm, forward, recon):
        """Load (or replace) one arm's forward and reconstruction COSY maps."""
        _check(library().simc_ctx_load_arm(self._h, arm, forward.encode(), recon.encode()))

    def close(self):
        if self._h:
            library().simc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- transport + reconstruction -------------------------------------------------
    def run(self, arm, events, *, p_spec, m2, ms=True, wcs=True, decay=False, ctau=0.0,
            coll=False, fit="ref_f32", wire_sigma=0.0, rng="philox", seed=0, event0=0,
            tape=None, stride=0, out=None, stream=None, stage_events=None, sieve=0):
        """Transport + reconstruct events through one arm.

        events: dict of float64 device tensors dpp, x, y, dxdz, dydz (+ z, fry).
        Returns a dict of device tensors (see include/simc_gpu.h for semantics):
        dpp, y, dxdz, dydz (recon or input), x_fp, dx_fp, y_fp, dy_fp, pathlen,
        resmult, m2, ok (uint8), stop (uint8), used (int32).
        stage_events: optional five recorded torch.cuda.Event(enable_timing=True)
        marking the boundaries of the three kernel stages and the
        reconstruction kernel (first launch chunk).
        sieve: the mock sieve slit (the reference's compile-time use_sieve in
        each mc_*.f), a bit mask: 0 off, 1 the arm's sieve (use_sieve), 2 the
        SHMS's second sieve in front of the horizontal bender (mc_shms.f
        use_2sieve), 3 both (SHMS)."""
        n = events["dpp"].numel()
        ev = {k: _dev_f64(events[k], n, k) for k in ("dpp", "x", "y", "dxdz", "dydz")}
        for k in ("z", "fry"):
            ev[k] = _dev_f64(events[k], n, k) if events.get(k) is not None else None
        dev = ev["dpp"].device
        if out is None:
            out = {k: ev[k].clone() for k in ("dpp", "y", "dxdz", "dydz")}
            for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "resmult", "m2"):
                out[k] = torch.zeros(n, dtype=torch.float64, device=dev)
            out["ok"] = torch.zeros(n, dtype=torch.uint8, device=dev)
            out["stop"] = torch.zeros(n, dtype=torch.uint8, device=dev)
            out["used"] = torch.zeros(n, dtype=torch.int32, device=dev)
        else:
            _check_outputs(out, n, RUN_REQUIRED)
        S = Settings(p_spec, m2, ctau, int(ms), int(wcs), int(decay), int(coll),
                     FIT_REF_F32 if fit == "ref_f32" else FIT_FP64, int(sieve), wire_sigma)
        ein = EventsIn(*[_ptr(ev.get(k)) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")])
        eout = EventsOut(*[_ptr(out.get(k)) for k in OUT_FIELDS])
        if rng == "tape":
            if tape is None:
                raise ValueError("tape rng needs tape and stride")
            _check_tape(tape, n, stride)
            R = Rng(RNG_TAPE, 0, 0, event0, _ptr(tape), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        if stage_events is None:
            _check(library().simc_arm_run(self._h, arm, C.byref(S), C.byref(ein), C.byref(eout), n,
                                          C.byref(R), C.c_void_p(st)))
        else:
            if len(stage_events) != 5 or any(e.cuda_event == 0 for e in stage_events):
                raise ValueError("stage_events: five events, each recorded once beforehand")
            evs = (_vp * 5)(*[e.cuda_event for e in stage_events])
            _check(library().simc_arm_run_timed(self._h, arm, C.byref(S), C.byref(ein), C.byref(eout),
                                                n, C.byref(R), C.c_void_p(st), evs))
        return out

    def run_host(self, arm, events, *, p_spec, m2, ms=True, wcs=True, decay=False, ctau=0.0,
                 coll=False, fit="ref_f32", wire_sigma=0.0, rng="philox", seed=0, event0=0,
                 tape=None, stride=0, out=None, sieve=0):
        """run() on numpy arrays through the synchronous host-pointer entry point.
        Returns a dict of numpy arrays with the same fields as run(); `out`
        (such a dict from an earlier call) is written in place instead (the
        INOUT columns x_fp..dy_fp and pathlen start from its values)."""
        n = len(events["dpp"])
        ev = {k: np.ascontiguousarray(events[k], dtype=np.float64) for k in
              ("dpp", "x", "y", "z", "dxdz", "dydz", "fry") if events.get(k) is not None}
        if out is None:
            out = {k: ev[k].copy() for k in ("dpp", "y", "dxdz", "dydz")}
            for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen", "resmult", "m2"):
                out[k] = np.zeros(n)
            out["ok"] = np.zeros(n, np.uint8)
            out["stop"] = np.zeros(n, np.uint8)
            out["used"] = np.zeros(n, np.int32)
        for k, v in out.items():
            if v is not None and not (isinstance(v, np.ndarray) and v.flags.c_contiguous and len(v) == n and
                                      v.dtype == {"ok": np.uint8, "stop": np.uint8, "used": np.int32}.get(k, np.float64)):
                raise TypeError(f"out[{k!r}]: expected a contiguous numpy array of {n} elements")
        S = Settings(p_spec, m2, ctau, int(ms), int(wcs), int(decay), int(coll),
                     FIT_REF_F32 if fit == "ref_f32" else FIT_FP64, int(sieve), wire_sigma)
        hp = lambda a: None if a is None else a.ctypes.data_as(_vp)
        ein = EventsIn(*[hp(ev.get(k)) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")])
        eout = EventsOut(*[hp(out.get(k)) for k in OUT_FIELDS])
        if rng == "tape":
            tape = np.ascontiguousarray(tape, dtype=np.float64)
            R = Rng(RNG_TAPE, 0, 0, event0, hp(tape), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        _check(library().simc_arm_run_host(self._h, arm, C.byref(S), C.byref(ein), C.byref(eout), n,
                                           C.byref(R)))
        return out

    def recon_coefficients(self, arm):
        """The loaded reconstruction map's coefficients, numpy [n_terms, 4]."""
        n = C.c_int32()
        _check(library().simc_recon_coefficients(self._h, arm, None, C.byref(n)))
        c = np.zeros((n.value, 4))
        _check(library().simc_recon_coefficients(self._h, arm, c.ctypes.data_as(C.POINTER(C.c_double)), None))
        return c

    def adjoint(self, arm, events, mask, coef, offsets, *, p_spec, m2, ms=True, wcs=True, wire_sigma=0.0,
                rng="philox", seed=0, event0=0, tape=None, stride=0, grad_coef=None, out9=None, stream=None):
        """Adjoint form (simc_adjoint_run): loss and gradients over the events
        with mask != 0, draws replayed, mask fixed.  coef: float64 device
        tensor [n_terms, 4]; offsets: 6 floats (ox, oy, ox', oy', s_p, dth).
        Returns device tensors (grad_coef [n_terms, 4], out9) where out9 =
        [loss, dL/d offsets (6), excluded, used]."""
        n = events["dpp"].numel()
        ev = {k: _dev_f64(events[k], n, k) for k in ("dpp", "x", "y", "dxdz", "dydz")}
        fry = events.get("fry")
        if fry is not None:
            ev["fry"] = _dev_f64(fry, n, "fry")
        if not (mask.is_cuda and mask.dtype == torch.uint8 and mask.numel() == n and mask.is_contiguous()):
            raise TypeError("mask: expected n contiguous uint8 on the device")
        if not (coef.is_cuda and coef.dtype == torch.float64 and coef.is_contiguous()):
            raise TypeError("coef: expected a contiguous float64 device tensor")
        if grad_coef is None:
            grad_coef = torch.empty_like(coef)
        if out9 is None:
            out9 = torch.empty(9, dtype=torch.float64, device=coef.device)
        S = Settings(p_spec, m2, 0.0, int(ms), int(wcs), 0, 0, FIT_FP64, 0, wire_sigma)
        ein = EventsIn(*[_ptr(ev.get(k)) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")])
        if rng == "tape":
            R = Rng(RNG_TAPE, 0, 0, event0, _ptr(tape), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        off = (C.c_double * 6)(*[float(v) for v in offsets])
        st = stream if stream is not None else torch.cuda.current_stream(coef.device).cuda_stream
        _check(library().simc_adjoint_run(self._h, arm, C.byref(S), C.byref(ein), _ptr(mask), n, C.byref(R),
                                          _ptr(coef), off, _ptr(grad_coef), _ptr(out9), C.c_void_p(st)))
        return grad_coef, out9

    # -- target step ---------------------------------------------------------------------
    def target_prep(self, target, prep, vertex, *, rng="philox", seed=0, event0=0, tape=None, stride=0,
                    out=None, stream=None):
        """Target step (simc_target_prep): vertex quantities (dict of float64
        device tensors, fields of simc_vertex_in; absent = NULL) to the arm's
        inputs.  target: Target; prep: PrepSettings.  Returns a dict of device
        tensors dpp, x, y, z, dxdz, dydz, fry, eloss, teff, dang1, dang2 and
        used (int32), or fills `out`."""
        n = vertex["z"].numel()
        vin = {k: _dev_f64(vertex[k], n, k) for k in VERTEX_FIELDS if vertex.get(k) is not None}
        dev = vin["z"].device
        if out is None:
            out = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in PREP_OUT_FIELDS if k != "used"}
            out["used"] = torch.empty(n, dtype=torch.int32, device=dev)
        else:
            for k in PREP_OUT_FIELDS:
                if out.get(k) is not None:
                    _dev_tensor(out[k], n, "out[" + k + "]", torch.int32 if k == "used" else torch.float64)
        V = VertexIn(*[_ptr(vin.get(k)) for k in VERTEX_FIELDS])
        O = PrepOut(*[_ptr(out.get(k)) for k in PREP_OUT_FIELDS])
        if rng == "tape":
            if tape is None or stride <= 0:
                raise ValueError("tape rng needs tape and stride")
            R = Rng(RNG_TAPE, 0, 0, event0, _ptr(tape), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        _check(library().simc_target_prep(self._h, C.byref(target), C.byref(prep), C.byref(V), C.byref(O), n,
                                          C.byref(R), C.c_void_p(st)))
        return out

    # -- event generator -----------------------------------------------------------------
    def gen_hep_rad(self, target, hep, rad, n, *, rng="philox", seed=0, event0=0, tape=None, stride=0, out=None,
                    device=None, stream=None):
        """H(e,e'p) elastic events with radiative tails (simc_gen_hep_rad):
        the VERTEX kinematics (HEP_OUT_FIELDS), the photons and ORIG
        kinematics (RAD_OUT_FIELDS), ntail (int32), ok (uint8), used
        (int32) as device tensors, or fills `out`."""
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if out is None:
            out = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in HEP_OUT_FIELDS + RAD_OUT_FIELDS}
            out["ntail"] = torch.empty(n, dtype=torch.int32, device=dev)
            out["ok"] = torch.empty(n, dtype=torch.uint8, device=dev)
            out["used"] = torch.empty(n, dtype=torch.int32, device=dev)
        for k in HEP_OUT_FIELDS + RAD_OUT_FIELDS:
            if out.get(k) is not None:
                _dev_tensor(out[k], n, k, torch.float64)
        for k, dt in (("ok", torch.uint8), ("used", torch.int32), ("ntail", torch.int32)):
            if out.get(k) is not None:
                _dev_tensor(out[k], n, k, dt)
        O = HepOut(*[_ptr(out.get(k)) for k in HEP_OUT_FIELDS + ("ok", "used")])
        RO = RadOut(*[_ptr(out.get(k)) for k in RAD_OUT_FIELDS + ("ntail",)])
        if rng == "tape":
            if tape is None:
                raise ValueError("tape rng needs tape and stride")
            _check_tape(tape, n, stride)
            R = Rng(RNG_TAPE, 0, 0, event0, _ptr(tape), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        _check(library().simc_gen_hep_rad(self._h, C.byref(target), C.byref(hep), C.byref(rad), C.byref(O),
                                          C.byref(RO), int(n), C.byref(R), C.c_void_p(st)))
        return out

    def gen_hep(self, target, hep, n, *, rng="philox", seed=0, event0=0, tape=None, stride=0, out=None,
                device=None, stream=None):
        """H(e,e'p) elastic events (simc_gen_hep): n vertices with their
        kinematics and weight.  target: Target; hep: HepSettings.  Returns a
        dict of float64 device tensors (HEP_OUT_FIELDS), ok (uint8) and used
        (int32), or fills `out` (absent keys are not written)."""
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if out is None:
            out = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in HEP_OUT_FIELDS}
            out["ok"] = torch.empty(n, dtype=torch.uint8, device=dev)
            out["used"] = torch.empty(n, dtype=torch.int32, device=dev)
        for k in HEP_OUT_FIELDS:
            if out.get(k) is not None:
                _dev_f64(out[k], n, k)
        for k, dt in (("ok", torch.uint8), ("used", torch.int32)):
            t = out.get(k)
            if t is not None and (not t.is_cuda or t.dtype != dt or t.numel() != n or not t.is_contiguous()):
                raise ValueError(f"{k}: expected {n} contiguous {dt} elements on a HIP device")
        O = HepOut(*[_ptr(out.get(k)) for k in HEP_OUT_FIELDS + ("ok", "used")])
        if rng == "tape":
            if tape is None or stride <= 0:
                raise ValueError("tape rng needs tape and stride")
            R = Rng(RNG_TAPE, 0, 0, event0, _ptr(tape), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        _check(library().simc_gen_hep(self._h, C.byref(target), C.byref(hep), C.byref(O), int(n), C.byref(R),
                                      C.c_void_p(st)))
        return out

    def gen_hep_host(self, target, hep, n, *, rng="philox", seed=0, event0=0, tape=None, stride=0):
        """simc_gen_hep_host: the generator into numpy arrays (synchronous)."""
        out = {k: np.zeros(n) for k in HEP_OUT_FIELDS}
        out["ok"] = np.zeros(n, np.uint8)
        out["used"] = np.zeros(n, np.int32)
        O = HepOut(*[C.c_void_p(out[k].ctypes.data) for k in HEP_OUT_FIELDS + ("ok", "used")])
        if rng == "tape":
            tape = np.ascontiguousarray(tape, dtype=np.float64)
            R = Rng(RNG_TAPE, 0, 0, event0, C.c_void_p(tape.ctypes.data), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        _check(library().simc_gen_hep_host(self._h, C.byref(target), C.byref(hep), C.byref(O), int(n), C.byref(R)))
        return out

    def target_prep_host(self, target, prep, vertex, *, rng="philox", seed=0, event0=0, tape=None, stride=0):
        """target_prep on numpy arrays (synchronous host-pointer entry point)."""
        n = len(vertex["z"])
        vin = {k: np.ascontiguousarray(vertex[k], dtype=np.float64) for k in VERTEX_FIELDS
               if vertex.get(k) is not None}
        out = {k: np.zeros(n) for k in PREP_OUT_FIELDS if k != "used"}
        out["used"] = np.zeros(n, np.int32)
        hp = lambda a: None if a is None else a.ctypes.data_as(_vp)
        V = VertexIn(*[hp(vin.get(k)) for k in VERTEX_FIELDS])
        O = PrepOut(*[hp(out.get(k)) for k in PREP_OUT_FIELDS])
        if rng == "tape":
            tape = np.ascontiguousarray(tape, dtype=np.float64)
            R = Rng(RNG_TAPE, 0, 0, event0, hp(tape), int(stride))
        else:
            R = Rng(RNG_PHILOX, 0, int(seed), int(event0), None, 0)
        _check(library().simc_target_prep_host(self._h, C.byref(target), C.byref(prep), C.byref(V), C.byref(O),
                                               n, C.byref(R)))
        return out

    def recon(self, arm, x_fp, dx_fp, y_fp, dy_fp, fry=None, stream=None):
        """mc_hms_recon / mc_shms_recon over focal-plane tracks."""
        n = x_fp.numel()
        ins = [_dev_f64(t, n, k) for t, k in ((x_fp, "x_fp"), (dx_fp, "dx_fp"), (y_fp, "y_fp"),
                                               (dy_fp, "dy_fp"))]
        fr = _dev_f64(fry, n, "fry") if fry is not None else None
        outs = [torch.empty_like(ins[0]) for _ in range(4)]
        st = stream if stream is not None else torch.cuda.current_stream(ins[0].device).cuda_stream
        _check(library().simc_recon_run(self._h, arm, n, *[_ptr(t) for t in ins], _ptr(fr),
                                        *[_ptr(t) for t in outs], C.c_void_p(st)))
        return tuple(outs)  # delta (%), x'_tar, y'_tar, y_tar (cm)

    # -- bookkeeping -------------------------------------------------------------------
    def counters(self, arm):
        names = COUNTERS[arm]
        buf = (C.c_int64 * len(names))()
        _check(library().simc_counters_get(self._h, arm, buf, len(names)))
        return dict(zip(names, list(buf)))

    def reset_counters(self, arm):
        _check(library().simc_counters_reset(self._h, arm))

    def class_info(self, arm, cls):
        d, a, n = C.c_double(), C.c_int32(), C.c_int32()
        _check(library().simc_map_class_info(self._h, arm, cls, C.byref(d), C.byref(a), C.byref(n)))
        return d.value, bool(a.value), n.value

    def poly_flops_full(self, arm):
        return int(library().simc_poly_flops_full(self._h, arm))

    def stop_flops(self, arm, stage=0):
        """Algorithmic polynomial FP64 operations of an event, by stop code
        (stage 0: the whole path; 1, 2: the first two kernel stages; 3: the
        reconstruction of an accepted event)."""
        buf = (C.c_int64 * 256)()
        _check(library().simc_stop_flops_stage(self._h, arm, stage, buf, 256))
        return list(buf)

    def hut_flops(self, arm, ms=True, wcs=True):
        """FP64 operations of the detector hut's steps for an event, by stop
        code (simc_stop_flops_hut)."""
        buf = (C.c_int64 * 256)()
        _check(library().simc_stop_flops_hut(self._h, arm, int(ms), int(wcs), buf, 256))
        return list(buf)


def _arm_call(ctx, arm, p_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag, decay_flag,
              fry, pathlen, using_coll, ctau, **kw):
    ev = {"dpp": dpp, "x": x, "y": y, "z": z, "dxdz": dxdz, "dydz": dydz, "fry": fry}
    n = dpp.numel()
    out = {"dpp": dpp, "y": y, "dxdz": dxdz, "dydz": dydz,
           "pathlen": pathlen if pathlen is not None else torch.zeros_like(dpp)}
    for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "resmult", "m2"):
        out[k] = torch.zeros_like(dpp)
    out["ok"] = torch.zeros(n, dtype=torch.uint8, device=dpp.device)
    out["stop"] = torch.zeros(n, dtype=torch.uint8, device=dpp.device)
    out["used"] = None
    ctx.run(arm, ev, p_spec=p_spec, m2=m2, ms=ms_flag, wcs=wcs_flag, decay=decay_flag,
            ctau=ctau, coll=using_coll, out=out, **kw)
    return out


def mc_hms(ctx, p_spec, th_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag, decay_flag,
           fry, pathlen=None, using_HMScoll=False, ctau=0.0, **kw):
    """Vectorised mc_hms (hms/mc_hms.f:1-4).  dpp, y, dxdz, dydz, pathlen are
    updated in place (INOUT); returns the output dict.  th_spec is unused by
    the reference too (hms/mc_hms.f:153)."""
    del th_spec
    return _arm_call(ctx, HMS, p_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag,
                     decay_flag, fry, pathlen, using_HMScoll, ctau, **kw)


def mc_shms(ctx, p_spec, th_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag, decay_flag,
            fry, pathlen=None, spectr=5, using_SHMScoll=False, ctau=0.0, **kw):
    """Vectorised mc_shms (shms/mc_shms.f:1-4); spectr must be 5 (:336-341)."""
    del th_spec
    if spectr != 5:
        raise SimcError(-6, "mc_shms: spectr must be 5 (the LSA tune is no longer supported)")
    return _arm_call(ctx, SHMS, p_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag,
                     decay_flag, fry, pathlen, using_SHMScoll, ctau, **kw)


def mc_hrsr(ctx, p_spec, th_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag, decay_flag,
            fry, pathlen=None, ctau=0.0, **kw):
    """Vectorised mc_hrsr (hrsr/mc_hrsr.f:1-3); y_fp of accepted events carries
    the VDC offset (hrsr/mc_hrsr.f:524)."""
    del th_spec
    return _arm_call(ctx, HRSR, p_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag,
                     decay_flag, fry, pathlen, False, ctau, **kw)


def mc_hrsl(ctx, p_spec, th_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag, decay_flag,
            fry, pathlen=None, ctau=0.0, **kw):
    """Vectorised mc_hrsl (hrsl/mc_hrsl.f:1-3); y_fp of accepted events carries
    the VDC offset (hrsl/mc_hrsl.f:525)."""
    del th_spec
    return _arm_call(ctx, HRSL, p_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag,
                     decay_flag, fry, pathlen, False, ctau, **kw)


def mc_sos(ctx, p_spec, th_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag, decay_flag,
           fry, pathlen=None, ctau=0.0, **kw):
    """Vectorised mc_sos (sos/mc_sos.f:1-3).  The SOS collimator is always the
    octagon pair (sos/mc_sos.f:173-204): there is no collimator switch."""
    del th_spec
    return _arm_call(ctx, SOS, p_spec, dpp, x, y, z, dxdz, dydz, m2, ms_flag, wcs_flag,
                     decay_flag, fry, pathlen, False, ctau, **kw)


FIELD_MAP = os.path.join(REPO, "data", "fields", "trg_field_map.dat")


class TargetField:
    """A polarized-target field on the device (simc_field_create = trgInit,
    trg_track.f:243-325).  map_path None = the uniform 5 T test field, "0" =
    no field; theta_e / theta_p: field angle relative to the electron / hadron
    arm in degrees."""

    def __init__(self, map_path=FIELD_MAP, theta_e=0.0, theta_p=0.0):
        h = _vp()
        _check(library().simc_field_create(None if map_path is None else map_path.encode(), theta_e, 0.0,
                                           theta_p, 0.0, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            library().simc_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def track_from_tgt(self, ev, mom, mass, spect, ok=None, stream=None):
        """track_from_tgt (trg_track.f:591-658) in place on the float64 device
        tensors ev["x"], ["y"], ["z"], ["dx"], ["dy"]; mom signed per event.
        Returns ok (uint8)."""
        n = ev["x"].numel()
        t = [_dev_f64(ev[k], n, k) for k in ("x", "y", "z", "dx", "dy")]
        m = _dev_f64(mom, n, "mom")
        ok = torch.zeros(n, dtype=torch.uint8, device=t[0].device) if ok is None else ok
        st = stream if stream is not None else torch.cuda.current_stream(t[0].device).cuda_stream
        _check(library().simc_track_from_tgt(self._h, spect, n, *[_ptr(v) for v in t], _ptr(m), mass, _ptr(ok),
                                             C.c_void_p(st)))
        return ok

    def track_to_tgt(self, ctx, arm, track, rec, frx, fry, mom, mass, ctheta, stheta, spect, ok=None,
                     stream=None):
        """track_to_tgt (trg_track.f:738-876) in place on rec["delta"], ["y"],
        ["dx"], ["dy"] (device tensors), with the arm's reconstruction of
        track["x_fp"], ["dx_fp"], ["y_fp"], ["dy_fp"] (the /track/ values).
        ok: uint8 device tensor INOUT (None: all true).  Returns ok."""
        n = rec["delta"].numel()
        tr = [_dev_f64(track[k], n, k) for k in ("x_fp", "dx_fp", "y_fp", "dy_fp")]
        io = [_dev_f64(rec[k], n, k) for k in ("delta", "y", "dx", "dy")]
        f = [_dev_f64(v, n, k) for v, k in ((frx, "frx"), (fry, "fry"), (mom, "mom"))]
        ok = torch.ones(n, dtype=torch.uint8, device=io[0].device) if ok is None else ok
        st = stream if stream is not None else torch.cuda.current_stream(io[0].device).cuda_stream
        _check(library().simc_track_to_tgt(ctx._h, self._h, arm, spect, n, *[_ptr(v) for v in tr],
                                           *[_ptr(v) for v in io], *[_ptr(v) for v in f], mass, ctheta, stheta,
                                           _ptr(ok), C.c_void_p(st)))
        return ok

    def track_from_tgt_host(self, ev, mom, mass, spect):
        """The same on numpy arrays (copies returned)."""
        import numpy as np
        out = {k: np.array(ev[k], dtype=np.float64, copy=True) for k in ("x", "y", "z", "dx", "dy")}
        n = len(out["x"])
        m = np.ascontiguousarray(mom, dtype=np.float64)
        ok = np.zeros(n, np.uint8)
        hp = lambda a: a.ctypes.data_as(_vp)  # noqa: E731
        _check(library().simc_track_from_tgt_host(self._h, spect, n, *[hp(out[k]) for k in ("x", "y", "z", "dx", "dy")],
                                                  hp(m), mass, hp(ok)))
        out["ok"] = ok
        return out


# Column names of the .bin ntuple per analysis mode, NtupleInit.f:36-341
# (the names are part of the format that util/root_tree and analyses read).
_NT_COMMON = ["hsdelta", "hsyptar", "hsxptar", "hsytar", "hsxfp", "hsxpfp", "hsyfp", "hsypfp", "hsdeltai",
              "hsyptari", "hsxptari", "hsytari", "ssdelta", "ssyptar", "ssxptar", "ssytar", "ssxfp", "ssxpfp",
              "ssyfp", "ssypfp", "ssdeltai", "ssyptari", "ssxptari", "ssytari", "q", "nu", "Q2", "W", "epsilon",
              "Em", "Pm", "thetapq", "phipq"]
_NT_TGT_FIELD = ["th_tarq", "phitarq", "beta", "phis", "phic", "betai", "phisi", "phici"]
_NT_PIZERO = ["xcal_gamma1", "ycal_gamma1", "Egamma1", "Pgamma1x", "Pgamma1y", "Pgamma1z", "xcal_gamma2",
              "ycal_gamma2", "Egamma2", "Pgamma2x", "Pgamma2y", "Pgamma2z"]


def ntuple_tags(mode, tgt_field=False, kaon=False, rho=False, pizero=False):
    """The ntuple's column names for mode 'elastic' (doing_hyd_elast,
    doing_deuterium, doing_heavy), 'pion' (doing_pion, doing_kaon,
    doing_delta) or 'semi' (doing_semi, doing_rho), NtupleInit.f:36-341."""
    t = list(_NT_COMMON)
    if mode == "pion":
        t += ["missmass", "mmnuc", "phad", "t", "pmpar", "pmper", "pmoop", "fry", "radphot", "pfermi", "siglab",
              "sigcm", "Weight", "decdist", "Mhadron", "pdotqhat", "Q2i", "Wi", "ti", "phipqi"]
        t += _NT_TGT_FIELD if tgt_field else []
        t += ["saghai", "factor"] if kaon else []
        t += _NT_PIZERO if pizero else []
    elif mode == "semi":
        t += ["missmass", "ppi", "t", "fry", "radphot", "siglab", "sigcent", "Weight", "decdist", "Mhadron", "z",
              "zi", "pt2", "pt2i", "xbj", "xbji", "thqi", "sighad", "jacobian", "centjac", "pfermi", "xfermi",
              "phipqi"]
        t += _NT_TGT_FIELD if tgt_field else []
        t += ["Mrho", "Thrho", "mmnuc"] if rho else []
        t += _NT_PIZERO if pizero else []
    elif mode == "elastic":
        t += ["corrsing", "Pmx", "Pmy", "Pmz", "PmPar", "PmPer", "PmOop", "fry", "radphot", "sigcc", "Weight",
              "theta_e", "theta_p"]
    else:
        raise ValueError("mode must be 'elastic', 'pion' or 'semi'")
    return t


class Ntuple:
    """The reference's .bin ntuple (simc_ntuple_*): open with the column
    names, append events from host rows or selected events of device
    columns, close."""

    def __init__(self, path, tags):
        self.tags = list(tags)
        arr = (C.c_char_p * len(self.tags))(*[t.encode() for t in self.tags])
        h = _vp()
        _check(library().simc_ntuple_open(path.encode(), len(self.tags), arr, C.byref(h)))
        self._h = h

    def write(self, rows):
        import numpy as np
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        if rows.ndim != 2 or rows.shape[1] != len(self.tags):
            raise ValueError(f"rows must be (n, {len(self.tags)})")
        _check(library().simc_ntuple_write(self._h, rows.ctypes.data_as(_vp), rows.shape[0]))

    def write_device(self, cols, select=None, stream=None):
        """cols: one float64 device tensor per column; select: uint8 device
        tensor (None: all events).  Returns the number of events written."""
        if len(cols) != len(self.tags):
            raise ValueError(f"expected {len(self.tags)} columns")
        n = cols[0].numel()
        cs = [_dev_f64(c, n, f"column {k}") for k, c in enumerate(cols)]
        ptrs = (_vp * len(cs))(*[c.data_ptr() for c in cs])
        if select is not None and (select.dtype != torch.uint8 or select.numel() != n or not select.is_cuda):
            raise TypeError("select: expected a uint8 device tensor of the columns' length")
        st = stream if stream is not None else torch.cuda.current_stream(cs[0].device).cuda_stream
        w = C.c_int64()
        _check(library().simc_ntuple_write_device(self._h, ptrs, _ptr(select), n, C.byref(w), C.c_void_p(st)))
        return w.value

    def close(self):
        if self._h:
            _check(library().simc_ntuple_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_ntuple(path):
    """(tags, rows) of a .bin ntuple: rows is an (events, columns) float64 array."""
    import numpy as np
    raw = open(path, "rb").read()
    pos = 0

    def record():
        nonlocal pos
        n = int.from_bytes(raw[pos:pos + 4], "little", signed=True)
        body = raw[pos + 4:pos + 4 + n]
        if int.from_bytes(raw[pos + 4 + n:pos + 8 + n], "little", signed=True) != n:
            raise ValueError(f"{path}: record framing broken at byte {pos}")
        pos += 8 + n
        return body
    ncols = int.from_bytes(record(), "little", signed=True)
    tags = [record().decode().rstrip() for _ in range(ncols)]
    rest = np.frombuffer(raw, dtype=np.uint8, offset=pos)
    if rest.size % (16 * ncols):
        raise ValueError(f"{path}: truncated event records")
    words = rest.view("<i4").reshape(-1, 4)
    if not ((words[:, 0] == 8) & (words[:, 3] == 8)).all():
        raise ValueError(f"{path}: event records are not 8-byte records")
    vals = rest.reshape(-1, 16)[:, 4:12].copy().view("<f8").reshape(-1, ncols)
    return tags, vals


def stop_report(ctx, path, electron_arm, hadron_arm):
    """The STOP-counter report of simc.f:462-530 (.geni) from ctx's counters."""
    _check(library().simc_stop_report(ctx._h, path.encode(), electron_arm, hadron_arm))


class Batch:
    """Many kinematic settings transported in one run (simc_batch_*).

    settings: list of dicts with keys arm, p_spec, m2, offset, count and
    optionally ms, wcs, decay, ctau, coll, fit, wire_sigma, sieve, seed, event0
    (the run() keywords).  Event j of a setting is element offset + j of the event
    arrays and draws Philox stream (seed, event0 + j)."""

    def __init__(self, ctx: Spectrometers, settings):
        arr = (Setting * len(settings))()
        for i, d in enumerate(settings):
            S = Settings(d["p_spec"], d["m2"], d.get("ctau", 0.0), int(d.get("ms", True)),
                         int(d.get("wcs", True)), int(d.get("decay", False)), int(d.get("coll", False)),
                         FIT_REF_F32 if d.get("fit", "ref_f32") == "ref_f32" else FIT_FP64, int(d.get("sieve", 0)),
                         d.get("wire_sigma", 0.0))
            arr[i] = Setting(d["arm"], 0, S, int(d.get("seed", 0)), int(d.get("event0", 0)),
                             int(d["offset"]), int(d["count"]))
        self.ctx = ctx
        self.settings = list(settings)
        self.n_arr = max(int(d["offset"]) + int(d["count"]) for d in settings)
        h = _vp()
        _check(library().simc_batch_create(ctx._h, arr, len(settings), C.byref(h)))
        self._h = h

    def _structs(self, events, out):
        n = self.n_arr
        ev = {k: _dev_tensor(events[k], n, k, torch.float64, at_least=True)
              for k in ("dpp", "x", "y", "dxdz", "dydz")}
        fry = events.get("fry")
        if fry is not None:
            ev["fry"] = _dev_tensor(fry, n, "fry", torch.float64, at_least=True)
        _check_outputs(out, n, RUN_REQUIRED, at_least=True)
        ein = EventsIn(*[_ptr(ev.get(k)) for k in ("dpp", "x", "y", "z", "dxdz", "dydz", "fry")])
        eout = EventsOut(*[_ptr(out.get(k)) for k in OUT_FIELDS])
        return ein, eout

    def new_outputs(self, events):
        dev = events["dpp"].device
        n = events["dpp"].numel()
        out = {k: events[k].clone() for k in ("dpp", "y", "dxdz", "dydz")}
        for k in ("x_fp", "dx_fp", "y_fp", "dy_fp", "pathlen"):
            out[k] = torch.zeros(n, dtype=torch.float64, device=dev)
        out["ok"] = torch.zeros(n, dtype=torch.uint8, device=dev)
        out["stop"] = torch.zeros(n, dtype=torch.uint8, device=dev)
        return out

    def run(self, events, out=None, stream=None):
        out = self.new_outputs(events) if out is None else out
        ein, eout = self._structs(events, out)
        st = stream if stream is not None else torch.cuda.current_stream(events["dpp"].device).cuda_stream
        _check(library().simc_batch_run(self._h, C.byref(ein), C.byref(eout), C.c_void_p(st)))
        return out

    def counters(self, i):
        names = COUNTERS[self.settings[i]["arm"]]
        buf = (C.c_int64 * len(names))()
        _check(library().simc_batch_counters(self._h, i, buf, len(names)))
        return dict(zip(names, list(buf)))

    def set_chunk(self, chunk):
        _check(library().simc_batch_set_chunk(self._h, int(chunk)))

    def passes(self):
        """(number of passes, passes whose lanes key their own setting's stream)."""
        n, k = C.c_int32(), C.c_int32()
        _check(library().simc_batch_passes(self._h, C.byref(n), C.byref(k)))
        return n.value, k.value

    def tune(self, events, out, stream=None):
        """Autotune the launch tile; returns (chunk, ms)."""
        ein, eout = self._structs(events, out)
        st = stream if stream is not None else torch.cuda.current_stream(events["dpp"].device).cuda_stream
        c, ms = C.c_int64(), C.c_double()
        _check(library().simc_batch_tune(self._h, C.byref(ein), C.byref(eout), C.c_void_p(st),
                                         C.byref(c), C.byref(ms)))
        return c.value, ms.value

    def close(self):
        if getattr(self, "_h", None):
            library().simc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
