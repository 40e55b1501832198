"""ctypes binding of libpirate_amd.so (C-ABI declared in include/pirate_amd.h).

The library is built in-tree by pirate_amd/build.py. There is no fallback: if the library
is missing, importing the dedisperser classes raises.
"""

import ctypes
import os

_here = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_here, 'libpirate_amd.so')

_lib = None

i64 = ctypes.c_int64
vp = ctypes.c_void_p


class TreeInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        'primary_tree_index', 'early_trigger_level', 'amb_rank', 'dd_rank', 'nt_ds',
        'max_width', 'dm_downsampling', 'time_downsampling', 'wt_dm_downsampling', 'wt_time_downsampling',
        'Dcore', 'nprofiles', 'ndm_out', 'ndm_wt', 'nt_out', 'nt_wt', 'nsubbands_N', 'nmultiplets_M', 'pf_rank')] + \
        [('dm_min', ctypes.c_double), ('dm_max', ctypes.c_double), ('trigger_frequency', ctypes.c_double)]


class GroupParams(ctypes.Structure):
    _fields_ = [('nslots', ctypes.c_int64), ('nbatches_out', ctypes.c_int64), ('device', ctypes.c_int32),
                ('use_graph', ctypes.c_int32)]


class DedisperserParams(ctypes.Structure):
    _fields_ = [('nbatches_out', ctypes.c_int64), ('nbatches_wt', ctypes.c_int64),
                ('num_consumers', ctypes.c_int64), ('device', ctypes.c_int32),
                ('synchronous', ctypes.c_int32), ('initial_chunk', ctypes.c_int64),
                ('enable_variances', ctypes.c_int32)]


class TileSource(ctypes.Structure):
    _fields_ = [('base', vp), ('dtype', ctypes.c_int32), ('nbeams', i64), ('beam_stride', i64), ('nrows', i64),
                ('row_stride', i64), ('t_lo', i64), ('t_hi', i64), ('origins', vp)]


# name -> (restype, argtypes)
_SIGS = {
    'pdd_last_error': (ctypes.c_char_p, []),
    'pdd_abi_version': (ctypes.c_int, []),
    'pdd_config_from_yaml_file': (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
    'pdd_config_from_yaml_string': (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
    'pdd_config_destroy': (None, [vp]),
    'pdd_config_to_yaml': (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    'pdd_config_make_channel_map': (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]),
    'pdd_config_eval': (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
    'pdd_config_test': (ctypes.c_int, [vp]),
    'pdd_plan_create': (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp)]),
    'pdd_plan_destroy': (None, [vp]),
    'pdd_plan_to_yaml': (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    'pdd_plan_ntrees': (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
    'pdd_plan_tree_info': (ctypes.c_int, [vp, i64, ctypes.POINTER(TreeInfo)]),
    'pdd_plan_decode_argmax': (ctypes.c_int, [vp, ctypes.c_uint32, i64, i64, i64, ctypes.POINTER(i64)]),
    'pdd_plan_decode_argmax_batch': (ctypes.c_int, [vp, i64, ctypes.POINTER(ctypes.c_uint32), i64, ctypes.POINTER(i64)]),
    'pdd_plan_decode_argmax2': (ctypes.c_int, [vp, i64, i64, i64, i64, i64, i64, ctypes.POINTER(ctypes.c_double)]),
    'pdd_plan_steady_state_it0': (ctypes.c_int, [vp, i64, ctypes.POINTER(i64), ctypes.c_size_t]),
    'pdd_plan_pf_variances': (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), i64, i64, ctypes.POINTER(ctypes.c_double), i64,
                                             ctypes.POINTER(ctypes.c_double), i64]),
    'pdd_dedisperser_create': (ctypes.c_int, [vp, ctypes.POINTER(DedisperserParams), ctypes.POINTER(vp)]),
    'pdd_dedisperser_destroy': (None, [vp]),
    'pdd_dedisperser_allocate': (ctypes.c_int, [vp]),
    'pdd_dedisperser_gpu_bytes': (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
    'pdd_dedisperser_host_bytes': (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
    'pdd_output_ipc_handle': (ctypes.c_int, [vp, vp, ctypes.POINTER(i64)]),
    'pdd_output_array_offset': (ctypes.c_int, [vp, i64, ctypes.c_int, ctypes.POINTER(i64)]),
    'pdd_ipc_open': (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp)]),
    'pdd_ipc_close': (ctypes.c_int, [vp, ctypes.c_int]),
    'pdd_acquire_input': (ctypes.c_int, [vp, i64, vp, ctypes.POINTER(vp)]),
    'pdd_release_input_and_launch': (ctypes.c_int, [vp, i64, vp]),
    'pdd_acquire_output': (ctypes.c_int, [vp, i64, i64, vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp), i64]),
    'pdd_release_output': (ctypes.c_int, [vp, i64, i64, vp]),
    'pdd_output_variances': (ctypes.c_int, [vp, i64, ctypes.POINTER(vp), i64]),
    'pdd_dequantize': (ctypes.c_int, [ctypes.c_char_p, i64, i64, i64, vp, vp, vp, vp]),
    'pdd_ringbuf_copy_create': (ctypes.c_int, [i64, i64, i64, ctypes.POINTER(ctypes.c_uint32), i64, ctypes.POINTER(vp)]),
    'pdd_ringbuf_copy_destroy': (None, [vp]),
    'pdd_ringbuf_copy_launch': (ctypes.c_int, [vp, vp, i64, i64, i64, vp]),
    'pdd_fill_all_weights': (ctypes.c_int, [vp, i64, ctypes.POINTER(ctypes.c_float), i64]),
    'pdd_fill_analytic_weights': (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), i64]),
    'pdd_stop': (ctypes.c_int, [vp, ctypes.c_char_p]),
    'pdd_kernel_timing': (ctypes.c_int, [vp, ctypes.c_int]),
    'pdd_kernel_timing_report': (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    'pdd_debug_copy': (ctypes.c_int, [vp, ctypes.c_char_p, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    'pdd_dedisperser_spill_bytes': (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
    'pdd_dedisperser_set_tuning': (ctypes.c_int, [vp, ctypes.c_char_p, i64]),
    'pdd_dedisperser_get_tuning': (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(i64)]),
    'pdd_group_create': (ctypes.c_int, [ctypes.POINTER(vp), i64, vp, ctypes.POINTER(vp)]),
    'pdd_group_destroy': (None, [vp]),
    'pdd_group_allocate': (ctypes.c_int, [vp]),
    'pdd_group_gpu_bytes': (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
    'pdd_group_member': (ctypes.c_int, [vp, i64, ctypes.POINTER(vp)]),
    'pdd_group_acquire_input': (ctypes.c_int, [vp, i64, vp, ctypes.POINTER(vp), i64]),
    'pdd_group_release_input_and_launch': (ctypes.c_int, [vp, i64, vp]),
    'pdd_group_acquire_output': (ctypes.c_int, [vp, i64, vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp), i64]),
    'pdd_group_release_output': (ctypes.c_int, [vp, i64, vp]),
    'pdd_group_stop': (ctypes.c_int, [vp, ctypes.c_char_p]),
    'pdd_group_set_use_graph': (ctypes.c_int, [vp, ctypes.c_int]),
    'pdd_group_set_tuning': (ctypes.c_int, [vp, i64, ctypes.c_char_p, i64]),
    'pdd_group_graph_nodes': (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
    'pdd_scorer_nparams': (i64, []),
    'pdd_scorer_forward': (ctypes.c_int, [vp, ctypes.POINTER(TileSource), i64, vp, vp]),
    'pdd_scorer_grad': (ctypes.c_int, [vp, ctypes.POINTER(TileSource), vp, i64, vp, vp, vp, vp]),
    'pdd_scorer_grad_noalloc': (ctypes.c_int, [vp, ctypes.POINTER(TileSource), vp, i64, vp, vp, vp, vp]),
    'pdd_scorer_adam': (ctypes.c_int, [vp, vp, vp, vp, i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, vp]),
    'pdd_dedisperser_tree_rows': (ctypes.c_int, [vp, i64, ctypes.POINTER(TileSource)]),
    'pdd_dedisperser_set_scorer': (ctypes.c_int, [vp, i64, ctypes.POINTER(ctypes.c_float), i64]),
    'pdd_output_scores': (ctypes.c_int, [vp, i64, ctypes.POINTER(vp), ctypes.POINTER(i64), ctypes.POINTER(i64)]),
}

EXPORTED_SYMBOLS = sorted(_SIGS)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f'pirate_amd: {LIBPATH} is not built; run "python -m pirate_amd.build" '
                               '(or __graft_entry__.build()) first')
        # torch wheels bundle their own HIP runtime with the same soname (libamdhip64.so.7)
        # as /opt/rocm's. Loading torch first makes this library bind to the runtime torch
        # uses, so the process holds one HIP runtime; loaded the other way round, torch
        # later finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PddError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise PddError(lib().pdd_last_error().decode(errors='replace'))


def get_string(fn, handle, *args):
    needed = ctypes.c_size_t(0)
    check(fn(handle, *args, None, 0, ctypes.byref(needed)))
    buf = ctypes.create_string_buffer(needed.value)
    check(fn(handle, *args, buf, needed.value, ctypes.byref(needed)))
    return buf.value.decode()
