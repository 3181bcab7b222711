"""ctypes binding of libarcwarden_hip.so (the C ABI declared in include/aw.h).

The library is loaded lazily by :func:`lib`.  If it has not been built, or the
HIP runtime cannot be loaded, :class:`LibraryMissing` is raised -- there is no
CPU fallback for any operation.  ``torch`` is imported first when available so
that the process keeps a single HIP runtime (torch's bundled copy and the
system copy share SONAMEs; whichever loads first serves both).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libarcwarden_hip.so"

AW_OK = 0
AW_ERR_INVALID, AW_ERR_HIP, AW_ERR_FFT, AW_ERR_NOMEM, AW_ERR_UNSUPPORTED, AW_ERR_STATE = \
    1, 2, 3, 4, 5, 6
AW_DEPOSIT_GLOBAL, AW_DEPOSIT_LDS_GRID, AW_DEPOSIT_BINNED, AW_DEPOSIT_AUTO = 0, 1, 2, 3
AW_CURRENT_CIC, AW_CURRENT_ESIRKEPOV = 0, 1
AW_J_STAGGERED = 1
AW_MODEL_ES, AW_MODEL_DARWIN = 0, 1


class LibraryMissing(RuntimeError):
    """libarcwarden_hip.so is not built or cannot be loaded."""


class AwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[aw status {code}] {msg}")
        self.code = code


class AwGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("dy", C.c_double),
                ("Lx", C.c_double), ("Ly", C.c_double)]


class AwRunParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("n0", C.c_double), ("vth", C.c_double), ("vd", C.c_double),
        ("wpe", C.c_double), ("wce", C.c_double), ("B0", C.c_float * 3), ("ppc", C.c_int32),
        ("nsteps", C.c_int64), ("dump_every", C.c_int64), ("norm", C.c_int32),
        ("eps0", C.c_double), ("qm", C.c_double), ("c", C.c_double), ("ndc", C.c_int32),
        ("jfilter", C.c_int32), ("tile_sort", C.c_int32), ("bnd_x", C.c_int32),
        ("bnd_nd", C.c_int32), ("bnd_numax", C.c_double), ("deltaf", C.c_int32),
        ("df_tpar", C.c_double), ("df_tperp", C.c_double), ("rel", C.c_int32),
        ("cold_nc", C.c_double), ("b0_prof", C.c_int32), ("b0_a", C.c_double),
        ("b0_xc", C.c_double), ("b0_lre", C.c_double), ("b0_sref", C.c_double),
        ("b0_c", C.c_double * 6), ("df_dist", C.c_int32), ("df_rho", C.c_double),
        ("df_kappa", C.c_double), ("df_taud", C.c_double), ("df_wprec", C.c_int32),
        ("ant_amp", C.c_double), ("ant_x0", C.c_double), ("ant_sigma", C.c_double),
        ("ant_w0", C.c_double), ("ant_trmp", C.c_double), ("ant_toff", C.c_double),
        ("pump", C.c_int32), ("pump_ex", C.c_double), ("pump_ey", C.c_double),
        ("pump_ez", C.c_double), ("pump_k0", C.c_double), ("pump_w0", C.c_double),
        ("pump_trmp", C.c_double), ("pump_toff", C.c_double), ("weight", C.c_double),
        ("ax", C.c_double), ("ay", C.c_double), ("k_filter", C.c_double),
        ("rng_seed", C.c_uint64), ("perturb_amp", C.c_double), ("perturb_kx", C.c_int32),
        ("two_stream", C.c_int32), ("bump_on_tail", C.c_int32), ("beam_frac", C.c_double),
        ("beam_vd", C.c_double), ("beam_vth", C.c_double), ("noisy_load", C.c_int32),
    ]


class AwSpecies(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("density", C.c_double), ("ppc", C.c_int32),
                ("uth", C.c_double * 3), ("ufl", C.c_double * 3), ("deltaf", C.c_int32),
                ("dist", C.c_int32), ("lc_rho", C.c_double), ("lc_kappa", C.c_double),
                ("taud", C.c_double), ("wdnoise", C.c_double)]


class AwYeeMesh(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz",
                                          "vcy", "vcz")]


class AwDeltafState(C.Structure):
    _fields_ = [("wd", C.c_void_p), ("wc", C.c_void_p), ("wdd", C.c_void_p)]


class AwCheckpointHeader(C.Structure):
    _fields_ = [("version", C.c_uint32), ("rng_seed", C.c_uint64), ("step", C.c_int64),
                ("time", C.c_double), ("sd", C.c_double), ("eps_l", C.c_double)]


class AwCheckpointArray(C.Structure):
    _fields_ = [("name", C.c_char_p), ("dtype", C.c_uint32), ("count", C.c_uint64),
                ("data", C.c_void_p), ("on_device", C.c_int32)]


class AwParticles(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("ux", C.c_void_p), ("uy", C.c_void_p),
                ("uz", C.c_void_p), ("w", C.c_void_p), ("n", C.c_int64),
                ("w_uniform", C.c_double)]


class AwFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "ELx", "ELy", "ETx", "ETy", "ETz")]


class AwSources(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("rho", "Jx", "Jy", "Jz", "dcux", "dcuy", "dcuz", "amu0", "amu1", "amu2",
                 "amu3")]


class AwDiagSample(C.Structure):
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("ke", C.c_double),
                ("ee", C.c_double), ("total", C.c_double), ("charge", C.c_double),
                ("max_e", C.c_double)]


class AwSimConfig(C.Structure):
    _fields_ = [("has_b0", C.c_int32), ("deposit_policy", C.c_int32),
                ("field_model", C.c_int32), ("tile_x", C.c_int32), ("tile_y", C.c_int32),
                ("sort_every", C.c_int32), ("stray_frac", C.c_double), ("n_modes", C.c_int32),
                ("mode_amp", C.c_double * 32), ("legacy_load", C.c_int32),
                ("fixed_scale", C.c_int32), ("blocks_per_tile", C.c_int32),
                ("autotune", C.c_int32), ("current_deposit", C.c_int32)]


class AwRunDesc(C.Structure):
    _fields_ = [("grid", AwGrid), ("rp", AwRunParams), ("species", C.c_void_p),
                ("nspecies", C.c_int32), ("has_b0", C.c_int32), ("n_modes", C.c_int32),
                ("mode_amp", C.c_double * 32)]


class AwEnsembleStats(C.Structure):
    _fields_ = [("nruns", C.c_int32), ("ngroups", C.c_int32), ("particles", C.c_int64),
                ("cells", C.c_int64), ("workgroups", C.c_int32), ("graph", C.c_int32),
                ("sorts", C.c_int64), ("graph_launches", C.c_int64), ("run_sorts", C.c_int64)]


AW_DECK_MAX_SPECIES = 16
AW_DECK_MAX_PROBES = 64


class AwDeck(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("Lx", C.c_double), ("Ly", C.c_double),
                ("rp", AwRunParams), ("nspecies", C.c_int32),
                ("species", AwSpecies * AW_DECK_MAX_SPECIES), ("dump_every", C.c_int64),
                ("outdir", C.c_char * 256), ("darwin", C.c_int32), ("yee", C.c_int32),
                ("dx_wpe_c", C.c_double), ("c_direct", C.c_double), ("wce", C.c_double),
                ("theta_deg", C.c_double), ("b0_direct", C.c_int32), ("b0_xc_set", C.c_int32),
                ("bnoise", C.c_double), ("pump_enable", C.c_int32), ("pump_M", C.c_int32),
                ("pump_w0", C.c_double), ("pump_ex0", C.c_double), ("pump_ey0", C.c_double),
                ("pump_ez0", C.c_double), ("pump_amp", C.c_double), ("pump_trmp", C.c_double),
                ("pump_toff", C.c_double), ("diag_enable", C.c_char * 256),
                ("nprobes", C.c_int32), ("probes", C.c_double * AW_DECK_MAX_PROBES),
                ("tsnap", C.c_double), ("band_lo", C.c_int32), ("band_hi", C.c_int32),
                ("kt_stride", C.c_int32), ("n_frames", C.c_int32), ("prefix", C.c_char * 128)]


P = C.c_void_p
PP = C.POINTER
_SIGS = {
    "aw_last_error": (None, [C.c_char_p, C.c_size_t]),
    "aw_version": (C.c_char_p, []),
    "aw_run_params_defaults": (None, [PP(AwRunParams)]),
    "aw_species_defaults": (None, [PP(AwSpecies)]),
    "aw_ctx_create": (C.c_int32, [PP(AwGrid), C.c_int32, C.c_int32, PP(P)]),
    "aw_ctx_destroy": (C.c_int32, [P]),
    "aw_ctx_set_fixed_scale": (C.c_int32, [P, C.c_int32]),
    "aw_ctx_tiles": (C.c_int32, [P, PP(C.c_int32), PP(C.c_int32), PP(C.c_int32),
                                 PP(C.c_int32)]),
    "aw_cell_index": (C.c_int32, [P, PP(AwParticles), P, P]),
    "aw_bin_tiles": (C.c_int32, [P, PP(AwParticles), P, P, P]),
    "aw_sort_cells": (C.c_int32, [P, PP(AwParticles), P, P, P]),
    "aw_sort_by_tile": (C.c_int32, [P, PP(AwParticles), PP(AwParticles), P, P]),
    "aw_migrate": (C.c_int32, [P, PP(AwParticles), P, P]),
    "aw_load_species": (C.c_int32, [P, PP(AwSpecies), C.c_int32, PP(AwRunParams),
                                    PP(C.c_double), C.c_int32, PP(AwParticles),
                                    PP(C.c_double), P]),
    "aw_load_legacy": (C.c_int32, [P, PP(AwRunParams), PP(AwParticles), P]),
    "aw_deposit_rho": (C.c_int32, [P, PP(AwParticles), C.c_double, C.c_int32, P, P]),
    "aw_deposit_rho_noalloc": (C.c_int32, [PP(AwGrid), PP(AwParticles), C.c_double, C.c_int32,
                                           C.c_int32, P, P]),
    "aw_rho_fixed_scale": (C.c_int32, [PP(AwGrid), C.c_double, C.c_double, C.c_double,
                                       C.c_int64, PP(C.c_int32)]),
    "aw_deposit_esirkepov": (C.c_int32, [P, P, P, PP(AwParticles), C.c_double, C.c_double,
                                         C.c_int32, P, P, P, P]),
    "aw_push_es": (C.c_int32, [P, PP(AwParticles), P, P, PP(AwRunParams), C.c_int32, P]),
    "aw_half_step_back": (C.c_int32, [P, PP(AwParticles), P, P, PP(AwRunParams), C.c_int32,
                                      P]),
    "aw_push_deposit_es": (C.c_int32, [P, PP(AwParticles), P, P, P, PP(AwRunParams),
                                       C.c_int32, P, PP(C.c_int64), P]),
    "aw_fft_r2c": (C.c_int32, [P, P, P, P]),
    "aw_fft_c2r": (C.c_int32, [P, P, P, P]),
    "aw_solve_es": (C.c_int32, [P, P, P, P, C.c_double, P]),
    "aw_solve_darwin_el_b": (C.c_int32, [P, P, P, C.c_double, C.c_double,
                                         PP(C.c_float), P]),
    "aw_solve_darwin_el_b_ex": (C.c_int32, [P, P, P, C.c_double, C.c_double,
                                            PP(C.c_float), C.c_int32, P]),
    "aw_solve_darwin_et": (C.c_int32, [P, P, P, C.c_double, C.c_double, C.c_double, P]),
    "aw_reduce_energies": (C.c_int32, [P, PP(AwParticles), P, P, P, C.c_double,
                                       PP(AwDiagSample), P]),
    "aw_sim_create": (C.c_int32, [PP(AwGrid), PP(AwRunParams), PP(AwSpecies), C.c_int32,
                                  PP(AwSimConfig), PP(P)]),
    "aw_sim_destroy": (C.c_int32, [P]),
    "aw_sim_init": (C.c_int32, [P]),
    "aw_sim_step": (C.c_int32, [P, C.c_int64]),
    "aw_sim_run": (C.c_int32, [P, C.c_int64]),
    "aw_sim_synchronize": (C.c_int32, [P]),
    "aw_sim_stream": (C.c_int32, [P, PP(P)]),
    "aw_sim_grid": (C.c_int32, [P, PP(AwGrid)]),
    "aw_sim_particles": (C.c_int32, [P, PP(AwParticles)]),
    "aw_sim_fields": (C.c_int32, [P, PP(AwFields)]),
    "aw_sim_sources": (C.c_int32, [P, PP(AwSources)]),
    "aw_sim_history": (C.c_int32, [P, PP(AwDiagSample), C.c_int64, PP(C.c_int64)]),
    "aw_sim_diag_now": (C.c_int32, [P, C.c_int64, PP(AwDiagSample)]),
    "aw_phase_hist": (C.c_int32, [PP(AwParticles), PP(AwGrid), C.c_int32, C.c_int32,
                                  C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_int32, P, P]),
    "aw_sum_sq": (C.c_int32, [P, C.c_int64, C.c_double, PP(C.c_double), P]),
    "aw_gather_f32": (C.c_int32, [P, P, C.c_int64, P, P]),
    "aw_device_synchronize": (C.c_int32, []),
    "aw_adj_push_reverse": (C.c_int32, [P, PP(AwParticles), P, P, PP(AwRunParams), C.c_int32,
                                        C.c_int32, P, P, P, P, P, P, P]),
    "aw_adj_gather_t": (C.c_int32, [P, PP(AwParticles), P, P, P, P, P]),
    "aw_adj_deposit_t": (C.c_int32, [P, PP(AwParticles), P, C.c_double, P, P, P]),
    "aw_checkpoint_write": (C.c_int32, [C.c_char_p, PP(AwCheckpointHeader), C.c_char_p,
                                        C.c_char_p, PP(AwCheckpointArray), C.c_int32]),
    "aw_checkpoint_open": (C.c_int32, [C.c_char_p, PP(P)]),
    "aw_checkpoint_close": (C.c_int32, [P]),
    "aw_checkpoint_info": (C.c_int32, [P, PP(AwCheckpointHeader), PP(C.c_int32),
                                       PP(C.c_int64), PP(C.c_int64)]),
    "aw_checkpoint_text": (C.c_int32, [P, C.c_int32, C.c_char_p, C.c_int64]),
    "aw_checkpoint_array_info": (C.c_int32, [P, C.c_int32, C.c_char_p, C.c_int64,
                                             PP(C.c_uint32), PP(C.c_uint64), PP(C.c_int32)]),
    "aw_checkpoint_read_array": (C.c_int32, [P, C.c_int32, P, C.c_int32]),
    "aw_sim_checkpoint_save": (C.c_int32, [P, C.c_char_p, C.c_int64, C.c_double, C.c_char_p,
                                           C.c_char_p]),
    "aw_sim_checkpoint_restore": (C.c_int32, [P, C.c_char_p]),
    "aw_maxwell_create": (C.c_int32, [PP(AwGrid), PP(AwRunParams), PP(P)]),
    "aw_maxwell_destroy": (C.c_int32, [P]),
    "aw_maxwell_stream": (C.c_int32, [P, PP(P)]),
    "aw_maxwell_synchronize": (C.c_int32, [P]),
    "aw_maxwell_mesh": (C.c_int32, [P, PP(AwYeeMesh)]),
    "aw_maxwell_step": (C.c_int32, [P, PP(AwParticles), PP(AwDeltafState)]),
    "aw_maxwell_step_at": (C.c_int32, [P, PP(AwParticles), PP(AwDeltafState), C.c_double]),
    "aw_maxwell_field_energy": (C.c_int32, [P, PP(C.c_double), PP(C.c_double)]),
    "aw_maxwell_residuals": (C.c_int32, [P, PP(AwParticles), P, PP(C.c_double),
                                         PP(C.c_double)]),
    "aw_maxwell_diag_fields": (C.c_int32, [P, C.c_double, PP(AwFields)]),
    "aw_wd_stats": (C.c_int32, [P, C.c_int64, PP(C.c_double), PP(C.c_double), PP(C.c_double),
                                P]),
    "aw_mirror_count": (C.c_int32, [PP(AwGrid), PP(AwRunParams), PP(AwSpecies),
                                    PP(C.c_int64)]),
    "aw_load_mirror": (C.c_int32, [PP(AwGrid), PP(AwRunParams), PP(AwSpecies),
                                   PP(AwParticles), P]),
    "aw_sim_request_sort": (C.c_int32, [P]),
    "aw_sim_particles_changed": (C.c_int32, [P]),
    "aw_sim_set_profiling": (C.c_int32, [P, C.c_int32]),
    "aw_sim_kernel_time": (C.c_int32, [P, PP(C.c_double), PP(C.c_int64), PP(C.c_int64)]),
    "aw_sim_exec_config": (C.c_int32, [P, PP(C.c_int32), PP(C.c_int32), PP(C.c_int32)]),
    "aw_deck_load": (C.c_int32, [C.c_char_p, PP(AwDeck)]),
    "aw_device_alloc": (C.c_int32, [C.c_size_t, PP(P)]),
    "aw_device_free": (C.c_int32, [P]),
    "aw_copy": (C.c_int32, [P, P, C.c_size_t]),
    "aw_memset_zero": (C.c_int32, [P, C.c_size_t, P]),
    "aw_stream_synchronize": (C.c_int32, [P]),
    "aw_ensemble_create": (C.c_int32, [PP(AwRunDesc), C.c_int32, PP(AwSimConfig), PP(P)]),
    "aw_ensemble_destroy": (C.c_int32, [P]),
    "aw_ensemble_set_graph": (C.c_int32, [P, C.c_int32]),
    "aw_ensemble_init": (C.c_int32, [P]),
    "aw_ensemble_step": (C.c_int32, [P, C.c_int64]),
    "aw_ensemble_run": (C.c_int32, [P, C.c_int64, C.c_int64]),
    "aw_ensemble_synchronize": (C.c_int32, [P]),
    "aw_ensemble_info": (C.c_int32, [P, PP(AwEnsembleStats)]),
    "aw_ensemble_particles": (C.c_int32, [P, C.c_int32, PP(AwParticles)]),
    "aw_ensemble_fields": (C.c_int32, [P, C.c_int32, PP(AwFields)]),
    "aw_ensemble_rho": (C.c_int32, [P, C.c_int32, PP(P)]),
    "aw_ensemble_history": (C.c_int32, [P, C.c_int32, PP(AwDiagSample), C.c_int64,
                                        PP(C.c_int64)]),
    "aw_adj_create": (C.c_int32, [PP(AwGrid), PP(AwRunParams), PP(AwSpecies), C.c_int32,
                                  C.c_int32, C.c_int32, C.c_int32, PP(P)]),
    "aw_adj_destroy": (C.c_int32, [P]),
    "aw_adj_forward": (C.c_int32, [P, PP(C.c_double), PP(C.c_double), C.c_int64,
                                   PP(C.c_double)]),
    "aw_adj_backward": (C.c_int32, [P, PP(C.c_double), PP(C.c_double)]),
    "aw_adj_state_grad": (C.c_int32, [P, P, P, P, P]),
    "aw_adj_initial_state": (C.c_int32, [P, PP(AwParticles)]),
    "aw_form_e": (C.c_int32, [P, PP(AwFields), PP(AwRunParams), C.c_double, P]),
    "aw_deposit_rho_j_amu": (C.c_int32, [P, PP(AwParticles), C.c_double, PP(AwSources), P]),
    "aw_deposit_dcu": (C.c_int32, [P, PP(AwParticles), C.c_double, PP(AwFields),
                                   PP(AwSources), P]),
    "aw_push_boris_em": (C.c_int32, [P, PP(AwParticles), PP(AwFields), PP(AwRunParams), P]),
}

_LIB = None


def exported_symbols() -> list[str]:
    """Names this binding expects the library to export (all of include/aw.h)."""
    return sorted(_SIGS)


def lib() -> C.CDLL:
    """Loads libarcwarden_hip.so once; raises LibraryMissing if unavailable."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run `python -m arcwarden_amd.build` first")
    try:  # one HIP runtime per process: let torch's copy load first if present
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        L = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    except OSError as e:
        raise LibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(2048)
    lib().aw_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != AW_OK:
        raise AwError(status, last_error())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def run_params(**kw) -> AwRunParams:
    rp = AwRunParams()
    lib().aw_run_params_defaults(C.byref(rp))
    for k, v in kw.items():
        if k == "B0":
            for i in range(3):
                rp.B0[i] = float(v[i])
        else:
            setattr(rp, k, v)
    return rp


def species(name="electrons", density=1.0, ppc=0, uth=(0, 0, 0), ufl=(0, 0, 0)) -> AwSpecies:
    sp = AwSpecies()
    lib().aw_species_defaults(C.byref(sp))
    sp.name = name.encode()[:31]
    sp.density = density
    sp.ppc = ppc
    for i in range(3):
        sp.uth[i] = uth[i]
        sp.ufl[i] = ufl[i]
    return sp


def grid(nx, ny, Lx, Ly) -> AwGrid:
    """Grid(nx, ny, Lx, Ly) as the reference's constructor (grid.hpp:53-56)."""
    return AwGrid(nx, ny, Lx / nx, Ly / ny, Lx, Ly)
