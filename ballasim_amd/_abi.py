"""ctypes mirror of include/bsx/bsx_sensor.h and the loader of the built library.

The product path is ballasim_amd/lib/libbsx.so (HIP, gfx950). load_library() raises if it is missing:
there is no fallback of any kind.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbsx.so")

BSX_OK, BSX_E_INVALID, BSX_E_LOGIC, BSX_E_RUNTIME = 0, 1, 2, 3
PHOTORECEPTOR = {"cascade": 0, "instantaneous": 1}
RESET = {"physical": 0, "ideal": 1}
PHASE = {"uniform": 0, "symmetric": 1, "zero": 2}
WAVE = {"constant": 0, "fourier": 1, "pulse": 2}


class Params(C.Structure):
    _fields_ = [
        ("family", C.c_char * 32),
        ("photoreceptor", C.c_int32),
        ("noise", C.c_int32),
        ("k", C.c_double), ("tau_o", C.c_double), ("tau_sf", C.c_double), ("sigma_o", C.c_double),
        ("illuminance_scale", C.c_double), ("veiling_glare", C.c_double),
        ("dark_current_fA", C.c_double), ("dark_sigma_ln", C.c_double), ("dark_ea_eV", C.c_double),
        ("temperature_K", C.c_double), ("reference_temperature_K", C.c_double),
        ("theta_on", C.c_double), ("theta_off", C.c_double), ("sigma_theta_on", C.c_double),
        ("sigma_theta_off", C.c_double), ("theta_onoff_corr", C.c_double), ("theta_floor", C.c_double),
        ("reset_mode", C.c_int32),
        ("initial_phase", C.c_int32),
        ("refractory_s", C.c_double), ("eps_ci", C.c_double),
        ("leak_rate_hz", C.c_double), ("leak_sigma_log10", C.c_double), ("leak_ea_eV", C.c_double),
        ("tau_hp", C.c_double), ("tau_hp_sigma_ln", C.c_double),
        ("hot_fraction", C.c_double),
        ("hot_rate_law", C.c_int32),
        ("limiter", C.c_int32),
        ("hot_rate_min_hz", C.c_double), ("hot_rate_max_hz", C.c_double), ("hot_rate_median_hz", C.c_double),
        ("hot_rate_sigma_ln", C.c_double), ("hot_leak_fraction", C.c_double),
        ("hot_reference_current_fA", C.c_double),
        ("readout_quantum_s", C.c_double), ("readout_delay_s", C.c_double),
        ("clock_offset_ns", C.c_int64),
        ("clock_drift_ppm", C.c_double),
        ("eps_L", C.c_double),
        ("noise_leaf_level", C.c_int32),
        ("noise_diffusion", C.c_int32),
        ("noise_margin_z", C.c_double),
        ("initial_state", C.c_int32),
        ("reserved0", C.c_int32),
        ("label_share_cut", C.c_double),
        ("shot_rate_hz", C.c_double),
    ]


class Waveform(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("n_harmonics", C.c_int32), ("f0_hz", C.c_double),
        ("a", C.c_double * 8), ("phi", C.c_double * 8),
        ("period_ns", C.c_double), ("phase_ns", C.c_double), ("rise_ns", C.c_double), ("top_ns", C.c_double),
        ("floor_level", C.c_double), ("peak_level", C.c_double),
    ]


class RawEvent(C.Structure):
    _fields_ = [("pixel", C.c_uint32), ("polarity", C.c_uint8), ("label", C.c_uint8), ("flags", C.c_uint16),
                ("t_cross", C.c_int64), ("ds", C.c_float), ("dn", C.c_float), ("dl", C.c_float),
                ("l_first", C.c_float)]


class DeviceRuns(C.Structure):
    _fields_ = [("pixel", C.c_void_p), ("row_start", C.c_void_p), ("row_count", C.c_void_p), ("n", C.c_size_t),
                ("t", C.c_void_p), ("irr", C.c_void_p), ("n_samples", C.c_size_t)]


class DeviceEvents(C.Structure):
    _fields_ = [("data", C.c_void_p), ("count", C.c_size_t)]


class Events(C.Structure):
    _fields_ = [("count", C.c_size_t)] + [
        (n, C.c_void_p) for n in
        ("x", "y", "p", "t_cross_ns", "t_reported_ns", "label", "flags", "ds", "dn", "dl", "l_first")]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "substeps", "substep_attempts", "settled_substeps", "cells", "refinements", "leaf_intervals", "fallbacks",
        "input_clamps", "quarantines", "pieces", "launches", "overflow_reruns", "events")] + [("kernel_ms", C.c_double)]


class LaunchPolicy(C.Structure):
    _fields_ = [("event_capacity", C.c_uint64), ("block_size", C.c_int32), ("walk_budget", C.c_int32),
                ("search_batch", C.c_int32), ("queue_chunk", C.c_int32), ("head_lanes", C.c_int32),
                ("head_per_mille", C.c_int32)]


class RigStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("graph_launches", C.c_uint64), ("captures", C.c_uint64),
                ("walk_reruns", C.c_uint64), ("kernel_ms", C.c_double)]


class Info(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("groups", C.c_int32), ("initialized", C.c_int32),
                ("time_ns", C.c_int64), ("readout_phase_ns", C.c_int64), ("irr_max", C.c_float),
                ("camera", C.c_uint32), ("seed", C.c_uint64)]


class Shape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("tex_seed", C.c_uint32), ("cx", C.c_double), ("cy", C.c_double),
                ("vx", C.c_double), ("vy", C.c_double), ("angle", C.c_double), ("omega", C.c_double),
                ("hx", C.c_double), ("hy", C.c_double), ("lux", C.c_float), ("contrast", C.c_float),
                ("octaves", C.c_int32), ("tex_cell_px", C.c_float)]


class View(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("origin_x", C.c_double), ("origin_y", C.c_double),
                ("px_scale", C.c_double)]


class SamplerRules(C.Structure):
    _fields_ = [("d_px", C.c_double), ("refine_tol_log", C.c_float), ("dark", C.c_float), ("l_min", C.c_int32),
                ("l_max", C.c_int32), ("max_refine_levels", C.c_int32), ("min_split_ns", C.c_uint32),
                ("refine_signature", C.c_int32), ("reserved", C.c_int32), ("fixed_period_ns", C.c_int64)]


P = C.c_void_p
# name -> (restype, argtypes); every entry point of include/bsx/*.h.
SIGNATURES = {
    "bsx_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "bsx_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "bsx_math_selfcheck": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "bsx_params_preset": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
    "bsx_preset_names": (C.c_int, [C.c_char_p, C.c_size_t]),
    "bsx_params_validate": (C.c_int, [C.POINTER(Params)]),
    "bsx_sensor_create": (C.c_int, [C.POINTER(Params), C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(Waveform),
                                    C.c_int32, C.c_uint32, C.POINTER(P)]),
    "bsx_sensor_destroy": (C.c_int, [P]),
    "bsx_sensor_set_photometry": (C.c_int, [P, C.c_double, P, C.c_size_t]),
    "bsx_sensor_init": (C.c_int, [P, C.c_int64, P]),
    "bsx_sensor_advance": (C.c_int, [P, P, C.c_size_t, P, P, C.c_size_t, P, C.c_int64]),
    "bsx_sensor_advance_runs": (C.c_int, [P, P, C.c_size_t, P, P, C.c_size_t, P, P, C.c_int64]),
    "bsx_sensor_advance_device": (C.c_int, [P, C.POINTER(DeviceRuns), C.c_int64, P, C.POINTER(DeviceEvents)]),
    "bsx_sensor_held_frame": (C.c_int, [P, P, P, P]),
    "bsx_sensor_advance_device_noalloc": (C.c_int, [P, C.POINTER(DeviceRuns), C.c_int64, P, P, C.c_uint64,
                                                    C.POINTER(C.c_uint64)]),
    "bsx_sensor_download_events": (C.c_int, [P, C.POINTER(DeviceEvents)]),
    "bsx_sensor_events": (C.c_int, [P, C.POINTER(Events)]),
    "bsx_sensor_state_bytes": (C.c_int, [P, C.POINTER(C.c_size_t)]),
    "bsx_sensor_save_state": (C.c_int, [P, P, C.c_size_t]),
    "bsx_sensor_load_state": (C.c_int, [P, P, C.c_size_t]),
    "bsx_sensor_stats": (C.c_int, [P, C.POINTER(Stats)]),
    "bsx_sensor_reset_stats": (C.c_int, [P]),
    "bsx_sensor_set_launch_policy": (C.c_int, [P, C.POINTER(LaunchPolicy)]),
    "bsx_sensor_info": (C.c_int, [P, C.POINTER(Info)]),
    "bsx_sensor_maps": (C.c_int, [P, P, P, P, P, P, P]),
    "bsx_rig_create": (C.c_int, [C.POINTER(P), C.c_int32, C.POINTER(P)]),
    "bsx_rig_destroy": (C.c_int, [P]),
    "bsx_rig_advance_device": (C.c_int, [P, C.POINTER(DeviceRuns), C.c_int64, P, C.POINTER(DeviceEvents)]),
    "bsx_rig_set_tuning": (C.c_int, [P, C.POINTER(LaunchPolicy)]),
    "bsx_rig_get_tuning": (C.c_int, [P, C.POINTER(LaunchPolicy)]),
    "bsx_rig_set_graph": (C.c_int, [P, C.c_int32]),
    "bsx_rig_stats": (C.c_int, [P, C.POINTER(RigStats)]),
    "bsx_rig_autotune": (C.c_int, [P, C.POINTER(DeviceRuns), C.c_int64, P, C.c_int32, C.POINTER(LaunchPolicy),
                                   C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32)]),
    "bsx_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(P)]),
    "bsx_mlp_destroy": (C.c_int, [P]),
    "bsx_mlp_set_params": (C.c_int, [P, P]),
    "bsx_mlp_get_params": (C.c_int, [P, P]),
    "bsx_mlp_gradients": (C.c_int, [P, P, P, C.c_int64, P, P, C.POINTER(C.c_double)]),
    "bsx_mlp_train": (C.c_int, [P, P, P, C.c_int64, C.c_int32, C.c_int32, C.c_float, P, C.POINTER(C.c_double)]),
    "bsx_mlp_predict": (C.c_int, [P, P, C.c_int64, P, P]),
    "bsx_mlp_tile_check": (C.c_int, [P, P, P]),
    "bsx_scene_learned_examples": (C.c_int, [P, C.POINTER(View), P, P, P, P, C.POINTER(C.c_uint64)]),
    "bsx_learned_sampler_create": (C.c_int, [P, C.POINTER(View), P, P, C.POINTER(P)]),
    "bsx_learned_sampler_destroy": (C.c_int, [P]),
    "bsx_learned_sampler_init": (C.c_int, [P, C.c_int64, P]),
    "bsx_learned_sampler_window": (C.c_int, [P, C.c_int64, C.c_int32, P, P, C.POINTER(DeviceRuns),
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "bsx_output_create": (C.c_int, [P, P, C.c_int64, C.POINTER(P)]),
    "bsx_output_destroy": (C.c_int, [P]),
    "bsx_output_window": (C.c_int, [P, C.POINTER(DeviceEvents), C.c_int64, P, P]),
    "bsx_output_flush": (C.c_int, [P, P, P]),
    "bsx_event_stream_create": (C.c_int, [C.c_size_t, P, P, C.POINTER(P)]),
    "bsx_event_stream_destroy": (C.c_int, [P]),
    "bsx_output_window_to_stream": (C.c_int, [P, C.POINTER(DeviceEvents), C.c_int64, P, P, P, P]),
    "bsx_output_flush_to_stream": (C.c_int, [P, P, P, P]),
    "bsx_event_stream_finish": (C.c_int, [P, P]),
    "bsx_event_stream_stats": (C.c_int, [P, P]),
    "bsx_scene_render_requests": (C.c_int, [P, C.POINTER(View), C.c_int64, P, C.c_size_t, P, C.c_int32, P, P, P, P]),
    "bsx_draw_maps": (C.c_int, [C.POINTER(Params), C.c_uint64, C.c_uint32, C.c_uint32, P, P, P, P, P]),
    "bsx_sensor_create_from_constants": (C.c_int, [P, P, C.c_size_t, P, C.c_size_t, C.POINTER(P)]),
    "bsx_sensor_upload_states": (C.c_int, [P, P, C.c_size_t]),
    "bsx_sensor_download_states": (C.c_int, [P, P, C.c_size_t]),
    "bsx_sensor_set_time": (C.c_int, [P, C.c_int64]),
    "bsx_sensor_set_kernel_photometry": (C.c_int, [P, C.c_float, C.c_float, P, C.c_size_t]),
    "bsx_sensor_raw_events": (C.c_int, [P, C.POINTER(P), C.POINTER(C.c_size_t)]),
    "bsx_noise_tables": (C.c_int, [C.POINTER(Params), P, P, C.c_size_t, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "bsx_scene_create": (C.c_int, [C.POINTER(Shape), C.c_int32, C.c_float, C.POINTER(P)]),
    "bsx_scene_destroy": (C.c_int, [P]),
    "bsx_copy_to_host": (C.c_int, [P, P, C.c_size_t]),
    "bsx_scene_render": (C.c_int, [P, C.POINTER(View), C.c_int64, P]),
    "bsx_scene_sample_counts": (C.c_int, [P, C.POINTER(View), P, C.c_int64, C.c_int32, P, C.POINTER(DeviceRuns),
                                          C.POINTER(C.c_uint64)]),
    "bsx_scene_set_capture_budget": (C.c_int, [P, C.c_uint64]),
    "bsx_scene_capture_budget": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "bsx_scene_sample_window": (C.c_int, [P, C.POINTER(View), C.POINTER(SamplerRules), C.c_int64, C.c_int32, P,
                                          C.POINTER(DeviceRuns), C.POINTER(C.c_uint64)]),
}

_lib = None


def _share_torch_runtime():
    """One HIP runtime per process. PyTorch-ROCm wheels bundle their own libamdhip64/libhsa-runtime64
    (soname libamdhip64.so.7, like /opt/rocm's); two copies in one process cannot share the device.
    When torch is installed, its runtime file is loaded first (without importing torch), so that
    libbsx binds to it and torch tensors can be handed to advance_device."""
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)


def load_library():
    """The built HIP library; raises if it is missing (build it with ballasim_amd.build.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"bsx: {LIB_PATH} is missing: the HIP library must be built "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    if os.environ.get("BSX_HIP_RUNTIME", "torch") == "torch":
        _share_torch_runtime()
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class BsxError(RuntimeError):
    pass


def check(status):
    """Raise the reference's exception type for a status code (ValueError for bad input)."""
    if status == BSX_OK:
        return
    buf = C.create_string_buffer(4096)
    load_library().bsx_last_error(buf, len(buf))
    msg = buf.value.decode(errors="replace")
    if status == BSX_E_INVALID:
        raise ValueError(msg)
    if status == BSX_E_LOGIC:
        raise BsxError(msg)
    raise BsxError(msg)
