"""ctypes binding of libgwfast_hip.so (the C ABI in include/gwfast_hip.h).

The library is built in-tree by ``gwfast_amd.build`` and loaded from
``gwfast_amd/lib/libgwfast_hip.so``.  There is no fallback: if the library or
a GPU is missing, ``lib()`` raises, and every public call of the package
fails loudly.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'lib', 'libgwfast_hip.so')

c_int32, c_int64, c_double, c_void_p, c_size_t = (ctypes.c_int32, ctypes.c_int64, ctypes.c_double,
                                                  ctypes.c_void_p, ctypes.c_size_t)
c_float = ctypes.c_float

MODEL_TAYLORF2 = 0
MODEL_IMRPHENOMD = 1
MODEL_IMRPHENOMD_NRTIDALV2 = 2
SHAPE_L = 0
SHAPE_T = 1
MAX_DETECTORS = 8


class GwfWaveform(ctypes.Structure):
    _fields_ = [('model', c_int32), ('is_tidal', c_int32), ('use_3p5PN_SpinHO', c_int32),
                ('phiref_vlso', c_int32), ('kerr_isco', c_int32), ('use_quadmon', c_int32),
                ('apply_fcut', c_int32), ('n_qnm', c_int32), ('fcut_par', c_double), ('fref', c_double),
                ('qnm_a', c_void_p), ('qnm_fring', c_void_p), ('qnm_fdamp', c_void_p),
                ('is_eccentric', c_int32), ('reserved', c_int32), ('fref_ecc', c_double)]


class GwfDetector(ctypes.Structure):
    _fields_ = [('shape', c_int32), ('earth_motion', c_int32), ('no_motion', c_int32), ('compute2arms', c_int32),
                ('lat_deg', c_double), ('long_deg', c_double), ('xax_deg', c_double), ('fmin', c_double),
                ('fmax', c_double), ('psd_f', c_void_p), ('psd_S', c_void_p), ('n_psd', c_int32),
                ('lin_spacing', c_int32), ('psd_f_first', c_double), ('psd_f_last', c_double)]


_P = ctypes.POINTER
# name -> (restype, argtypes), mirroring include/gwfast_hip.h
SIGNATURES = {
    'gwf_abi_version': (ctypes.c_int, []),
    'gwf_last_error': (ctypes.c_char_p, []),
    'gwf_workspace_bytes': (c_size_t, [c_int64]),
    'gwf_num_arms': (ctypes.c_int, [_P(GwfDetector), c_int32]),
    'gwf_fcut': (ctypes.c_int, [_P(GwfWaveform), c_void_p, c_int64, c_int64, c_double, c_void_p, c_void_p]),
    'gwf_grid': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_void_p, c_int64, c_int64, c_int32, c_void_p,
                                c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_waveform_eval': (ctypes.c_int, [_P(GwfWaveform), c_void_p, c_int64, c_int64, c_void_p, c_int64, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_snr': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64, c_int32,
                               c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_fisher': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64, c_int32,
                                  c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_fisher_noalloc': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64,
                                          c_int32, c_int32, c_void_p, c_void_p]),
    'gwf_snr_ragged': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64,
                                      c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_fisher_ragged': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64,
                                         c_void_p, c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_prepare': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64, c_int32,
                                   c_void_p, c_size_t, c_void_p]),
    'gwf_snr_prepared': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_int64, c_int32, c_void_p,
                                        c_void_p, c_size_t, c_void_p]),
    'gwf_fisher_prepared': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_int64, c_int32, c_int32,
                                           c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_strain': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64, c_void_p,
                                  c_int64, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_strain_components': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64,
                                             c_void_p, c_int64, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_grid_fcut': (ctypes.c_int, [_P(GwfDetector), c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p]),
    'gwf_overlap': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int32, c_void_p,
                                   c_void_p]),
    'gwf_strain_vjp': (ctypes.c_int, [c_void_p, c_void_p, c_int32, c_int64, c_int64, c_double, c_void_p, c_void_p]),
    'gwf_gram': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int64, c_int64, c_int32, c_void_p,
                                c_void_p]),
    'gwf_gram_split': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int64, c_int64, c_int32, c_void_p,
                                      c_void_p]),
    'gwf_cov_invert': (ctypes.c_int, [c_void_p, c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p]),
    'gwf_inversion_error': (ctypes.c_int, [c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_void_p]),
    'gwf_check_fisher': (ctypes.c_int, [c_void_p, c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    'gwf_fisher_reparam': (ctypes.c_int, [c_void_p, c_int64, c_void_p, c_int32, c_int64, c_int32, c_int32, c_void_p]),
    'gwf_cov_invert_ex': (ctypes.c_int, [c_void_p, c_int32, c_int64, c_int32, c_int32, c_int32, c_double, c_double,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'gwf_fisher_masked': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64,
                                         c_int32, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_set_queue_spin_limit': (ctypes.c_int, [ctypes.c_uint32]),
    'gwf_fisher_arms': (ctypes.c_int, [_P(GwfWaveform), _P(GwfDetector), c_int32, c_void_p, c_int64, c_int64, c_int32,
                                       c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_fisher_block_transform': (ctypes.c_int, [c_void_p, c_void_p, c_int32, c_int64, c_int32, c_int32, c_void_p,
                                                  c_void_p, c_void_p]),
    'gwf_fix_params': (ctypes.c_int, [c_void_p, c_int32, c_int64, ctypes.c_uint32, c_void_p, c_void_p, c_void_p]),
    'gwf_sigma_sky_area': (ctypes.c_int, [c_void_p, c_int32, c_int64, c_void_p, c_int32, c_int32, c_double, c_int32,
                                          c_void_p, c_void_p, c_void_p]),
    # learned SNR screen
    'gwf_gemm_bf16': (ctypes.c_int, [c_int32, c_int32, c_int32, c_float, c_void_p, c_int64, c_int64, c_void_p,
                                     c_int64, c_int64, c_float, c_void_p, c_int64, c_int32, c_void_p]),
    'gwf_bias_silu': (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p]),
    'gwf_silu_backward': (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    'gwf_colsum': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_int32, c_void_p]),
    'gwf_head_forward': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    'gwf_head_backward': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                         c_void_p]),
    'gwf_mse_grad': (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    'gwf_adam': (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_float, c_float, c_float, c_float,
                                c_int32, c_void_p, c_void_p]),
    'gwf_to_bf16': (ctypes.c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    'gwf_snr_features': (ctypes.c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    'gwf_snr_screen': (ctypes.c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                                      c_float, c_void_p, c_void_p, c_void_p]),
    'gwf_select_workspace_bytes': (c_size_t, [c_int64]),
    'gwf_select': (ctypes.c_int, [c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    'gwf_gather_events': (ctypes.c_int, [c_void_p, c_int64, c_int32, c_void_p, c_int64, c_void_p, c_int64,
                                         c_void_p]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the shared library and bind every ABI symbol (no GPU needed)."""
    if not os.path.exists(path):
        raise NativeError('libgwfast_hip.so not found at %s: run __graft_entry__.build() '
                          '(or python -m gwfast_amd.build) first' % path)
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.gwf_abi_version() != 1:
        raise NativeError('ABI version mismatch')
    return L


def lib():
    """The bound library; raises when it is missing or no GPU is visible."""
    global _lib
    if _lib is None:
        import torch
        if not torch.cuda.is_available():
            raise NativeError('gwfast_amd needs a ROCm GPU: torch.cuda.is_available() is False')
        _lib = load()
    return _lib


def check(rc, L=None):
    if rc != 0:
        L = L or _lib
        msg = L.gwf_last_error().decode() if L is not None else ''
        if rc == -22:
            raise ValueError(msg)
        raise NativeError('gwfast_hip error %d: %s' % (rc, msg))


def ptr(t):
    """Device pointer of a contiguous torch tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def as_f64_rows(ev_rows):
    return np.ascontiguousarray(np.asarray(ev_rows, dtype=np.float64))
