"""flint_amd -- MI355X implementation of FLINT's fft_small multiplication path.

Python view of the C ABI declared in include/flint_amd.h (libflint_amd.so,
built in-tree by flint_amd/build.py).  The functions mirror FLINT's
operator interface for the path (names and argument meaning follow
src/fft_small.h:850-865); data is passed as numpy uint64 arrays for the
host-pointer entry points and as raw device addresses for the _device
variants.  There is no CPU fallback: every call runs on the GPU and an
error is raised when the library or the device is unavailable.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflint_amd.so")

_u64 = ctypes.c_uint64
_s64 = ctypes.c_int64
_P64 = ctypes.POINTER(ctypes.c_uint64)
_PD = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p

FA_OK = 0


class FlintAmdError(RuntimeError):
    pass


class NmodT(ctypes.Structure):
    """FLINT's nmod_t (flint.h.in:441-447)."""
    _fields_ = [("n", _u64), ("ninv", _u64), ("norm", _u64)]


class Request(ctypes.Structure):
    """fa_request (include/flint_amd.h): one product of a batch."""
    _fields_ = [("kind", ctypes.c_int32), ("status", ctypes.c_int32),
                ("a", _VP), ("b", _VP), ("an", _s64), ("bn", _s64), ("aw", _s64), ("bw", _s64),
                ("z", _VP), ("zl", _s64), ("zh", _s64), ("zw", _s64), ("mod", NmodT),
                ("abits", _s64), ("bbits", _s64), ("np", ctypes.c_int32), ("bits", ctypes.c_int32),
                ("depth", ctypes.c_int32), ("trunc", ctypes.c_int32)]


REQ_MPN, REQ_NMOD, REQ_FMPZ = 0, 1, 2
TRUNC_AUTO, TRUNC_OFF, TRUNC_ON = 0, 1, 2
BATCH_DEVICE, BATCH_GRAPH, BATCH_FMPZ_KS = 1, 2, 4


class SdFftCtx(ctypes.Structure):
    """fa_sd_fft_ctx_struct (include/flint_amd.h)."""
    _fields_ = [("p", ctypes.c_double), ("pinv", ctypes.c_double), ("mod", NmodT),
                ("primitive_2power_root", _u64), ("depth", ctypes.c_int32),
                ("slot", ctypes.c_int32)]


class StreamReq(ctypes.Structure):
    """fa_stream_req (include/flint_amd.h): one header of the full stream."""
    _fields_ = [("kind", ctypes.c_int32), ("bits", ctypes.c_int32), ("an", _s64), ("bn", _s64),
                ("n", _u64), ("key", _u64 * 2)]


class FillSpec(ctypes.Structure):
    """fa_fill_spec (include/flint_amd.h): one operand written on the device."""
    _fields_ = [("dst", _VP), ("count", _u64), ("key", _u64), ("n", _u64), ("w", _u64),
                ("kind", ctypes.c_int32), ("bits", ctypes.c_int32)]


class EvalSpec(ctypes.Structure):
    """fa_eval_spec (include/flint_amd.h): one evaluation of device words."""
    _fields_ = [("src", _VP), ("count", _u64), ("w", _u64), ("x", _u64), ("m", _u64),
                ("is_signed", ctypes.c_int32), ("pad", ctypes.c_int32)]


class MpnCtxStruct(ctypes.Structure):
    """fa_mpn_ctx_struct (include/flint_amd.h)."""
    _fields_ = [("p", _u64), ("primes", _u64 * 8), ("device", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("state", _VP), ("magic", _u64)]


class PlanStruct(ctypes.Structure):
    """fa_plan_struct (include/flint_amd.h)."""
    _fields_ = [("kind", ctypes.c_int32), ("np", ctypes.c_int32), ("offset", ctypes.c_int32),
                ("bits", ctypes.c_int32), ("depth", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("stride", _u64), ("zl", _u64), ("zh", _u64), ("zn", _u64),
                ("an_max", _u64), ("bn_max", _u64), ("len_bound", _u64), ("mod", NmodT)]


class OpStruct(ctypes.Structure):
    """fa_op_struct (include/flint_amd.h)."""
    _fields_ = [("data", _VP), ("stride", _u64), ("np", _u64), ("offset", _u64),
                ("depth", _u64), ("domain", ctypes.c_int32), ("owns_data", ctypes.c_int32)]


OP_PRIMAL, OP_PRODUCT, OP_COEFF = 0, 1, 2

# C types of the drop-ins' CPU fallbacks (fa_set_cpu_fallback)
MPN_MUL_FN = ctypes.CFUNCTYPE(None, _P64, _P64, _s64, _P64, _s64)
NMOD_MUL_MID_FN = ctypes.CFUNCTYPE(None, _P64, _s64, _s64, _P64, _s64, _P64, _s64, NmodT)

# exported symbols and their signatures: (restype, argtypes)
SIGNATURES = {
    "mpn_mul_default_mpn_ctx": (None, [_P64, _P64, _s64, _P64, _s64]),
    "_nmod_poly_mul_mid_default_mpn_ctx": (None, [_P64, _s64, _s64, _P64, _s64, _P64, _s64, NmodT]),
    "fa_mpn_mul": (ctypes.c_int, [_P64, _P64, _s64, _P64, _s64]),
    "fa_mpn_mul_device": (ctypes.c_int, [_VP, _VP, _s64, _VP, _s64, _VP]),
    "fa_mpn_mul_range": (ctypes.c_int, [_P64, _u64, _u64, _P64, _s64, _P64, _s64]),
    "fa_mpn_mul_device_noalloc": (ctypes.c_int, [_VP, _VP, _s64, _VP, _s64, _VP]),
    "fa_noalloc_scratch_bytes": (_u64, []),
    "fa_mpn_mulmod_2expp1": (ctypes.c_int, [_P64, _P64, _P64, _s64]),
    "fa_mpn_mulmod_2expp1_device": (ctypes.c_int, [_VP, _VP, _VP, _s64, _VP]),
    "fa_mpn_mul_range_device": (ctypes.c_int, [_VP, _u64, _u64, _VP, _s64, _VP, _s64, _VP]),
    "fa_mpn_mul_range_threads": (ctypes.c_int, [_VP, _u64, _u64, _VP, _s64, _VP, _s64,
                                                ctypes.c_int, _VP]),
    "fa_mpn_mul_range_host": (ctypes.c_int, [_P64, _u64, _u64, _P64, _s64, _P64, _s64,
                                             ctypes.c_int]),
    "fa_mpn_window_plan": (ctypes.c_int, [_u64, _u64, _u64, ctypes.c_int, _P64]),
    "fa_mpn_plan": (ctypes.c_int, [_s64, _s64, _P64]),
    "fa_mpn_mul_plan": (ctypes.c_int, [_P64, _P64, _s64, _P64, _s64, ctypes.c_int, ctypes.c_int]),
    "fa_mpn_mul_device_plan": (ctypes.c_int, [_VP, _VP, _s64, _VP, _s64, ctypes.c_int,
                                              ctypes.c_int, _VP]),
    "fa_nmod_poly_mul_mid": (ctypes.c_int, [_P64, _s64, _s64, _P64, _s64, _P64, _s64, NmodT]),
    "fa_nmod_poly_mul_mid_device": (ctypes.c_int, [_VP, _s64, _s64, _VP, _s64, _VP, _s64, NmodT,
                                                   _VP]),
    "fa_nmod_poly_inv_series": (ctypes.c_int, [_P64, _P64, _s64, _s64, NmodT]),
    "fa_fmpz_poly_mul_mid_device": (ctypes.c_int, [_VP, _u64, _s64, _s64, _VP, _u64, _s64, _VP,
                                                   _u64, _s64, _s64, _s64, ctypes.c_int, _VP]),
    "fa_nmod_poly_mulmid_repack": (ctypes.c_int, [_P64, _P64, _s64, _P64, _s64, _s64, _s64,
                                                  NmodT]),
    "fa_nmod_poly_mulmid_fft_small": (ctypes.c_int, [_P64, _P64, _s64, _P64, _s64, _s64, _s64,
                                                     NmodT]),
    "fa_nmod_poly_divrem": (ctypes.c_int, [_P64, _P64, _P64, _s64, _P64, _s64, NmodT]),
    "fa_nmod_poly_divrem_precomp_init": (ctypes.c_int, [ctypes.POINTER(_VP), _P64, _s64, _s64,
                                                        NmodT]),
    "fa_nmod_poly_divrem_precomp": (ctypes.c_int, [_P64, _P64, _P64, _s64, _VP, NmodT]),
    "fa_nmod_poly_divrem_precomp_clear": (None, [_VP]),
    "fa_fmpz_poly_mul_mid": (ctypes.c_int, [_P64, _u64, _s64, _s64, _P64, _u64, _s64, _P64, _u64,
                                            _s64]),
    "fa_sd_fft_ctx_init_prime": (ctypes.c_int, [ctypes.POINTER(SdFftCtx), _u64]),
    "fa_sd_fft_ctx_clear": (None, [ctypes.POINTER(SdFftCtx)]),
    "fa_sd_fft_trunc": (ctypes.c_int, [ctypes.POINTER(SdFftCtx), _PD, _u64, _u64, _u64]),
    "fa_sd_ifft_trunc": (ctypes.c_int, [ctypes.POINTER(SdFftCtx), _PD, _u64, _u64]),
    "fa_sd_fft_ctx_w": (ctypes.c_double, [ctypes.POINTER(SdFftCtx), _u64]),
    "fa_last_error": (ctypes.c_char_p, []),
    "fa_init": (ctypes.c_int, []),
    "fa_default_primes": (ctypes.c_int, [_P64]),
    "fa_sync": (ctypes.c_int, []),
    "fa_stream": (_VP, []),
    "fa_synth_limbs": (None, [_u64, _P64, _u64, ctypes.c_int]),
    "fa_synth_mpn_pair": (None, [_u64, _P64, _u64, _P64, _u64]),
    "fa_rng_seed": (None, [_VP, _u64]),
    "fa_rng_next": (_u64, [_VP]),
    "fa_rng_uniform": (_u64, [_VP, _u64, _u64]),
    "fa_rng_double": (ctypes.c_double, [_VP]),
    "fa_rng_log_uniform": (_u64, [_VP, ctypes.c_double, ctypes.c_double]),
    "fa_rng_random_prime": (_u64, [_VP, ctypes.c_int]),
    "fa_rng_fill_below": (None, [_VP, _VP, _u64, _u64]),
    "fa_rng_fill_limbs": (None, [_VP, _VP, _u64, ctypes.c_int]),
    "fa_rng_fill_signed": (None, [_VP, _VP, _u64, ctypes.c_int, _u64]),
    "fa_stream_headers": (ctypes.c_int, [_u64, _s64, _VP]),
    "fa_stream_word": (_u64, [_u64, _u64]),
    "fa_stream_fill_device": (ctypes.c_int, [_VP, _s64, _VP]),
    "fa_eval_device": (ctypes.c_int, [_VP, _s64, _VP, _VP]),
    "fa_set_truncation": (None, [ctypes.c_int]),
    "fa_prof_enable": (None, [ctypes.c_int]),
    "fa_prof_reset": (None, []),
    "fa_prof_read": (ctypes.c_int, [ctypes.c_char_p, _PD, _P64, ctypes.c_int]),
    "fa_copy_device": (ctypes.c_int, [_VP, _VP, _u64, _VP]),
    "fa_set_fast_kernels": (None, [ctypes.c_int]),
    "fa_set_col_variant": (None, [ctypes.c_int]),
    "fa_set_row_variant": (None, [ctypes.c_int]),
    "fa_set_mpn_block_cap": (None, [ctypes.c_uint64]),
    "fa_set_poly_block_cap": (None, [ctypes.c_uint64]),
    "fa_set_ks_variant": (None, [ctypes.c_int]),
    "fa_set_prime_groups": (None, [ctypes.c_int]),
    "fa_set_crt_variant": (None, [ctypes.c_int]),
    "fa_set_square_route": (None, [ctypes.c_int]),
    "fa_plan_init_mpn": (ctypes.c_int, [_VP, _u64, _u64, _u64, ctypes.c_int]),
    "fa_plan_init_mpn_np": (ctypes.c_int, [_VP, _u64, _u64, _u64, ctypes.c_int, _u64]),
    "fa_plan_init_nmod_cyclic": (ctypes.c_int, [_VP, _u64, _u64, _u64, _u64, NmodT]),
    "fa_plan_init_mpn_cyclic": (ctypes.c_int, [_VP, _P64, _u64]),
    "fa_nmod_poly_mul_precomp_init": (ctypes.c_int, [ctypes.POINTER(_VP), _P64, _s64, _s64,
                                                     ctypes.c_int, NmodT]),
    "fa_nmod_poly_mul_mid_precomp": (ctypes.c_int, [_P64, _u64, _u64, _P64, _s64, _VP, NmodT]),
    "fa_nmod_poly_mul_precomp_clear": (None, [_VP]),
    "fa_mulmod_2expp1_ctx_init": (ctypes.c_int, [_VP, _s64, _s64]),
    "fa_mulmod_satisfies_bounds": (ctypes.c_int, [_u64]),
    "fa_nmod_direct_count": (_u64, []),
    "fa_set_direct_fft": (None, [ctypes.c_int]),
    "fa_sd_fft_ctx_point_mul": (ctypes.c_int, [_VP, _VP, _VP, _u64, _u64]),
    "fa_sd_fft_ctx_point_sqr": (ctypes.c_int, [_VP, _VP, _u64, _u64]),
    "fa_mulmod_2expp1_ctx_clear": (None, [_VP]),
    "fa_mulmod_2expp1": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fa_mulmod_2expp1_transformed_size": (_s64, [_VP]),
    "fa_mulmod_2expp1_transform": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "fa_mulmod_2expp1_mul_cached": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fa_plan_init_nmod": (ctypes.c_int, [_VP, _u64, _u64, _u64, _u64, _u64, _u64, NmodT]),
    "fa_plan_clear": (None, [_VP]),
    "fa_op_init": (ctypes.c_int, [_VP, _VP]),
    "fa_op_init_borrowed": (ctypes.c_int, [_VP, _VP, _VP]),
    "fa_op_sizeof_data": (_u64, [_VP]),
    "fa_op_clear": (None, [_VP]),
    "fa_fft_mpn": (ctypes.c_int, [_VP, _VP, _s64, _VP, _VP]),
    "fa_fft_nmod": (ctypes.c_int, [_VP, _VP, _s64, NmodT, _VP, _VP]),
    "fa_op_mul": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fa_op_sqr": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "fa_op_addmul": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fa_op_submul": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fa_op_add": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fa_op_sub": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fa_op_neg": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "fa_ifft": (ctypes.c_int, [_VP, _VP, _VP]),
    "fa_export_mpn": (ctypes.c_int, [_VP, _s64, _VP, _VP, _VP]),
    "fa_export_mpn_signed": (ctypes.c_int, [_VP, _s64, _VP, _VP, _u64, _VP, _VP]),
    "fa_export_mpn_signed_trunc": (ctypes.c_int, [_VP, _s64, _VP, _VP, _u64, _u64, _VP, _VP]),
    "fa_export_nmod": (ctypes.c_int, [_VP, _VP, NmodT, _VP, _VP]),
    "fa_export_nmod_range": (ctypes.c_int, [_VP, _VP, _u64, _u64, NmodT, _VP, _VP]),
    "fa_set_device_memory_limit": (None, [_u64]),
    "fa_set_cpu_fallback": (None, [_VP, _VP]),
    "fa_last_status": (ctypes.c_int, []),
    "fa_mpn_ctx_init": (ctypes.c_int, [_VP, _u64]),
    "fa_mpn_ctx_clear": (None, [_VP]),
    "fa_get_default_mpn_ctx": (_VP, []),
    "fa_mpn_ctx_mpn_mul": (ctypes.c_int, [_VP, _P64, _P64, _s64, _P64, _s64]),
    "fa_mpn_ctx_mpn_mul_device": (ctypes.c_int, [_VP, _VP, _VP, _s64, _VP, _s64, _VP]),
    "fa_nmod_poly_mul_mid_mpn_ctx": (ctypes.c_int, [_P64, _s64, _s64, _P64, _s64, _P64, _s64,
                                                    NmodT, _VP]),
    "fa_fmpz_poly_mul_mid_mpn_ctx": (ctypes.c_int, [_P64, _u64, _s64, _s64, _P64, _u64, _s64,
                                                    _P64, _u64, _s64, _VP]),
    "fa_batch_create": (ctypes.c_int, [ctypes.POINTER(_VP), _VP, _s64, ctypes.c_int]),
    "fa_batch_launch": (ctypes.c_int, [_VP, _VP]),
    "fa_batch_destroy": (None, [_VP]),
    "fa_batch_mul": (ctypes.c_int, [_VP, _s64, ctypes.c_int]),
    "fa_mpn_candidates": (ctypes.c_int, [_s64, _s64, _VP]),
    "fa_plan_features": (ctypes.c_int, [_s64, _s64, _VP]),
    "fa_autotune_mpn": (ctypes.c_int, [_s64, _s64, ctypes.c_int, _VP]),
    "fa_mlp_create": (ctypes.c_int, [ctypes.POINTER(_VP), _u64]),
    "fa_mlp_destroy": (None, [_VP]),
    "fa_mlp_param_count": (ctypes.c_int, []),
    "fa_mlp_get_params": (ctypes.c_int, [_VP, _VP]),
    "fa_mlp_set_params": (ctypes.c_int, [_VP, _VP]),
    "fa_mlp_forward": (ctypes.c_int, [_VP, _VP, ctypes.c_int, _VP]),
    "fa_mlp_loss_grad": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int, _VP, _VP]),
    "fa_mlp_train": (ctypes.c_int, [_VP, _VP, _VP, _VP, _s64, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, _VP]),
    "fa_mlp_plan_mpn": (ctypes.c_int, [_VP, _VP, ctypes.c_int, _VP]),
    "fa_fmpz_vec_max_bits": (_s64, [_VP, _u64, _s64]),
    "fa_fmpz_poly_mul_ks": (ctypes.c_int, [_P64, _u64, _s64, _s64, _P64, _u64, _s64, _P64, _u64,
                                           _s64]),
}

_lib = None


def lib():
    """Load libflint_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlintAmdError("libflint_amd.so is not built (run flint_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise FlintAmdError("flint_amd error %d: %s" % (rc, lib().fa_last_error().decode()))
    return rc


def _p(a):
    return a.ctypes.data_as(_P64)


def _arr(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def default_primes():
    out = np.zeros(8, dtype=np.uint64)
    _check(lib().fa_default_primes(_p(out)))
    return [int(x) for x in out]


def synth_limbs(seed, n, top_nonzero=True):
    a = np.empty(n, dtype=np.uint64)
    lib().fa_synth_limbs(seed, _p(a), n, 1 if top_nonzero else 0)
    return a


def synth_mpn_pair(seed, an, bn):
    a = np.empty(an, dtype=np.uint64)
    b = np.empty(bn, dtype=np.uint64)
    lib().fa_synth_mpn_pair(seed, _p(a), an, _p(b), bn)
    return a, b


def prof_enable(on=True):
    lib().fa_prof_enable(1 if on else 0)


def prof_reset():
    lib().fa_prof_reset()


def prof_read(max_entries=64):
    """{kernel name: (total ms, launches)} recorded since prof_reset()."""
    names = ctypes.create_string_buffer(64 * max_entries)
    ms = np.zeros(max_entries, dtype=np.float64)
    cnt = np.zeros(max_entries, dtype=np.int64)
    n = lib().fa_prof_read(names, ms.ctypes.data_as(_PD), cnt.ctypes.data_as(_P64), max_entries)
    raw = names.raw
    out = {}
    for i in range(n):
        nm = raw[64 * i: 64 * i + 64].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[i]), int(cnt[i]))
    return out


def set_fast_kernels(on=True):
    lib().fa_set_fast_kernels(1 if on else 0)


def set_row_variant(v):
    """row pass of the two-level transforms: 1 = three workgroups per CU,
    last-stage twiddles formed per layer (default); 0 = two, all held"""
    lib().fa_set_row_variant(int(v))


def set_prime_groups(v):
    """2^24-point products: primes in v groups on forked streams (default 2);
    0 or 1 = one launch per pass over all primes"""
    lib().fa_set_prime_groups(int(v))


def set_ks_variant(v):
    """Kronecker route for fmpz products beyond eight primes: 1 = signed slots,
    one integer product (default); 0 = three products of non-negative parts"""
    lib().fa_set_ks_variant(int(v))


def set_mpn_block_cap(limbs):
    """integer products above `limbs` total limbs go in device-added block
    products (0: only those no single plan covers)"""
    lib().fa_set_mpn_block_cap(ctypes.c_uint64(int(limbs)))


def set_poly_block_cap(n):
    """nmod/fmpz products above n output coefficients go in device-added
    block products (0: only above one 2^24-point transform)"""
    lib().fa_set_poly_block_cap(ctypes.c_uint64(int(n)))


def set_col_variant(v):
    """column-pass tiles at 2^24 points: 1 = 8 columns per workgroup, 0 = 4"""
    lib().fa_set_col_variant(int(v))


def set_truncation(mode=1):
    """row-granularity truncation of 2^13..2^23-point transforms (outputs
    identical): 0 never, 1 automatic (default), 2 wherever it applies"""
    lib().fa_set_truncation(int(mode))


def set_square_route(on=True):
    """squares: one forward transform per prime (default) or the general product"""
    lib().fa_set_square_route(1 if on else 0)


def set_crt_variant(v):
    """CRT of integer products: 1 = FP64 radix-2^24 digits, 0 = 64-bit limbs"""
    lib().fa_set_crt_variant(int(v))


def copy_device(dst_ptr, src_ptr, nbytes, stream=None):
    return _check(lib().fa_copy_device(dst_ptr, src_ptr, nbytes, stream))


def mpn_plan(an, bn):
    out = np.zeros(6, dtype=np.int64)
    _check(lib().fa_mpn_plan(an, bn, out.ctypes.data_as(_P64)))
    return dict(zip(("np", "bits", "depth", "d1", "d2", "zlen"), (int(x) for x in out)))


def mpn_mul(a, b):
    """mpn_mul_default_mpn_ctx: full product of two limb arrays."""
    a = _arr(a)
    b = _arr(b)
    z = np.empty(len(a) + len(b), dtype=np.uint64)
    _check(lib().fa_mpn_mul(_p(z), _p(a), len(a), _p(b), len(b)))
    return z


def mpn_mul_range(a, b, lo, hi, nthreads=None):
    """limbs [lo, hi) of a*b as _mpn_ctx_mpn_mul_range returns them
    (flint_mpn_mulmid_fft_small) with FLINT on one thread, or on nthreads
    threads (0: the exact window)"""
    a = _arr(a)
    b = _arr(b)
    z = np.empty(hi - lo, dtype=np.uint64)
    if nthreads is None:
        _check(lib().fa_mpn_mul_range(_p(z), lo, hi, _p(a), len(a), _p(b), len(b)))
    else:
        _check(lib().fa_mpn_mul_range_host(_p(z), lo, hi, _p(a), len(a), _p(b), len(b),
                                           nthreads))
    return z


def mpn_mul_range_device(z_ptr, lo, hi, a_ptr, an, b_ptr, bn, stream=None):
    return _check(lib().fa_mpn_mul_range_device(z_ptr, lo, hi, a_ptr, an, b_ptr, bn, stream))


def mpn_mul_range_threads(z_ptr, lo, hi, a_ptr, an, b_ptr, bn, nthreads, stream=None):
    """device window with the words of FLINT on nthreads threads (0: exact)"""
    return _check(lib().fa_mpn_mul_range_threads(z_ptr, lo, hi, a_ptr, an, b_ptr, bn, nthreads,
                                                 stream))


def mpn_window_plan(an, bn, lo, nthreads=1):
    """(two_prime, bits, k0, nk): the reference's plan as the window sees it"""
    out = np.zeros(4, dtype=np.int64)
    _check(lib().fa_mpn_window_plan(an, bn, lo, nthreads, out.ctypes.data_as(_P64)))
    return tuple(int(x) for x in out)


def mpn_mulmod_2expp1(x, y, n):
    """x*y mod 2^(64 n) + 1 for (n+1)-limb x, y <= 2^(64 n) (sd_fft_mpn_mulmod_2expp1)"""
    x = _arr(x)
    y = x if y is x else _arr(y)
    z = np.empty(n + 1, dtype=np.uint64)
    _check(lib().fa_mpn_mulmod_2expp1(_p(z), _p(x), _p(y), n))
    return z


def mpn_mul_plan(a, b, np_, bits):
    """product with a caller-chosen plan (np primes, bits-bit chunks)"""
    a = _arr(a)
    b = _arr(b)
    z = np.empty(len(a) + len(b), dtype=np.uint64)
    _check(lib().fa_mpn_mul_plan(_p(z), _p(a), len(a), _p(b), len(b), np_, bits))
    return z


def mpn_mul_device_plan(z_ptr, a_ptr, an, b_ptr, bn, np_, bits, stream=None):
    return _check(lib().fa_mpn_mul_device_plan(z_ptr, a_ptr, an, b_ptr, bn, np_, bits, stream))


def mpn_mul_device(z_ptr, a_ptr, an, b_ptr, bn, stream=None):
    """Device-resident product: raw device addresses, enqueued on `stream`."""
    return _check(lib().fa_mpn_mul_device(z_ptr, a_ptr, an, b_ptr, bn, stream))


def set_device_memory_limit(nbytes):
    """cap on the device bytes the library's grow-only buffers may hold
    (0: none); growth past it fails with FA_E_NOMEM"""
    lib().fa_set_device_memory_limit(ctypes.c_uint64(int(nbytes)))


def last_status():
    """status of this thread's last drop-in call (0, or the FA_E_* code of
    the device failure that sent it to the CPU fallback)"""
    return int(lib().fa_last_status())


class MpnCtx:
    """mpn_ctx_t (fft_small.h:382-433): a context with its own HIP stream and
    device buffers; MpnCtx.default() is the thread's (get_default_mpn_ctx)."""

    def __init__(self, p=0x0003F00000000001, _handle=None):
        if _handle is not None:
            self.ptr = _handle
            self._own = None
            return
        self._own = MpnCtxStruct()
        self.ptr = ctypes.addressof(self._own)
        _check(lib().fa_mpn_ctx_init(self.ptr, p))

    @classmethod
    def default(cls):
        return cls(_handle=lib().fa_get_default_mpn_ctx())

    def close(self):
        if self._own is not None:
            lib().fa_mpn_ctx_clear(self.ptr)
            self._own = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mpn_mul(self, a, b):
        """mpn_ctx_mpn_mul"""
        a = _arr(a)
        b = _arr(b)
        z = np.empty(len(a) + len(b), dtype=np.uint64)
        _check(lib().fa_mpn_ctx_mpn_mul(self.ptr, _p(z), _p(a), len(a), _p(b), len(b)))
        return z

    def mpn_mul_device(self, z_ptr, a_ptr, an, b_ptr, bn, stream=None):
        return _check(lib().fa_mpn_ctx_mpn_mul_device(self.ptr, z_ptr, a_ptr, an, b_ptr, bn,
                                                      stream))

    def nmod_poly_mul_mid(self, a, b, zl, zh, n):
        """_nmod_poly_mul_mid_mpn_ctx"""
        a = _arr(a)
        b = _arr(b)
        z = np.zeros(max(zh - zl, 1), dtype=np.uint64)
        _check(lib().fa_nmod_poly_mul_mid_mpn_ctx(_p(z), zl, zh, _p(a), len(a), _p(b), len(b),
                                                  nmod_init(n), self.ptr))
        return z[: max(zh - zl, 0)]

    def fmpz_poly_mul_mid(self, a, aw, b, bw, zl, zh, zw):
        """_fmpz_poly_mul_mid_mpn_ctx; None where the reference declines"""
        a = _arr(a)
        b = _arr(b)
        z = np.zeros(max((zh - zl) * zw, 1), dtype=np.uint64)
        rc = _check(lib().fa_fmpz_poly_mul_mid_mpn_ctx(_p(z), zw, zl, zh, _p(a), aw, len(a) // aw,
                                                       _p(b), bw, len(b) // bw, self.ptr))
        return z[: (zh - zl) * zw] if rc == 1 else None


class Plan:
    """fft_small_plan_t on the device path: Plan.mpn(an_max, bn_max, len_bound)
    or Plan.nmod(zl, zh, zn, xtrunc_max, len_bound, prod_bits, n)."""

    def __init__(self):
        self.s = PlanStruct()
        self.ptr = ctypes.addressof(self.s)

    @classmethod
    def mpn(cls, an_max, bn_max, len_bound=1, signed=False, np=None):
        """np: None or 0 the automatic choice, 2 two primes, else three or more
        (fft_small_plan_init_mpn_np)"""
        P = cls()
        if np is None:
            _check(lib().fa_plan_init_mpn(P.ptr, an_max, bn_max, len_bound, 1 if signed else 0))
        else:
            _check(lib().fa_plan_init_mpn_np(P.ptr, an_max, bn_max, len_bound,
                                             1 if signed else 0, np))
        return P

    @classmethod
    def mpn_cyclic(cls, nn, len_bound=1):
        """products mod 2^(64 nn) - 1; returns (plan, the rounded-up nn)"""
        P = cls()
        v = ctypes.c_uint64(nn)
        _check(lib().fa_plan_init_mpn_cyclic(P.ptr, ctypes.byref(v), len_bound))
        return P, int(v.value)

    @classmethod
    def nmod_cyclic(cls, depth, zh, len_bound, prod_bits, n):
        """convolutions mod x^(2^depth) - 1 over Z/nZ, coefficients [0, zh)"""
        P = cls()
        _check(lib().fa_plan_init_nmod_cyclic(P.ptr, depth, zh, len_bound, prod_bits,
                                              nmod_init(n)))
        return P

    @classmethod
    def nmod(cls, zl, zh, zn, xtrunc_max, len_bound, prod_bits, n):
        P = cls()
        _check(lib().fa_plan_init_nmod(P.ptr, zl, zh, zn, xtrunc_max, len_bound, prod_bits,
                                       nmod_init(n)))
        return P


class Mulmod2expp1Ctx(ctypes.Structure):
    """fa_mulmod_2expp1_ctx_struct: products mod 2^N + 1 with cached
    transforms (sd_fft_mpn_mulmod_2expp1_ctx)."""
    _fields_ = [("N", ctypes.c_int64), ("n", ctypes.c_int64), ("m", ctypes.c_int64),
                ("P", PlanStruct)]

    def __init__(self, N, m=0):
        super().__init__()
        _check(lib().fa_mulmod_2expp1_ctx_init(ctypes.addressof(self), N, m))

    @property
    def ptr(self):
        return ctypes.addressof(self)

    def transformed_size(self):
        return lib().fa_mulmod_2expp1_transformed_size(self.ptr)

    def transform(self, F_ptr, y_ptr, stream=None):
        return _check(lib().fa_mulmod_2expp1_transform(self.ptr, F_ptr, y_ptr, stream))

    def mul_cached(self, z_ptr, x_ptr, F_ptr, stream=None):
        return _check(lib().fa_mulmod_2expp1_mul_cached(self.ptr, z_ptr, x_ptr, F_ptr, stream))

    def mul(self, z_ptr, x_ptr, y_ptr, stream=None):
        return _check(lib().fa_mulmod_2expp1(self.ptr, z_ptr, x_ptr, y_ptr, stream))


class Op:
    """fft_small_op_t: a transformed operand in device memory."""

    def __init__(self, plan):
        self.s = OpStruct()
        self.ptr = ctypes.addressof(self.s)
        self.plan = plan
        _check(lib().fa_op_init(self.ptr, plan.ptr))

    def close(self):
        if self.s.data:
            lib().fa_op_clear(self.ptr)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def domain(self):
        return self.s.domain

    def fft_mpn(self, a_ptr, an, stream=None):
        return _check(lib().fa_fft_mpn(self.ptr, a_ptr, an, self.plan.ptr, stream))

    def fft_nmod(self, a_ptr, an, n, stream=None):
        return _check(lib().fa_fft_nmod(self.ptr, a_ptr, an, nmod_init(n), self.plan.ptr, stream))

    def mul(self, A, B, stream=None):
        return _check(lib().fa_op_mul(self.ptr, A.ptr, B.ptr, self.plan.ptr, stream))

    def sqr(self, A, stream=None):
        return _check(lib().fa_op_sqr(self.ptr, A.ptr, self.plan.ptr, stream))

    def addmul(self, A, B, stream=None):
        return _check(lib().fa_op_addmul(self.ptr, A.ptr, B.ptr, self.plan.ptr, stream))

    def submul(self, A, B, stream=None):
        return _check(lib().fa_op_submul(self.ptr, A.ptr, B.ptr, self.plan.ptr, stream))

    def add(self, A, B, stream=None):
        return _check(lib().fa_op_add(self.ptr, A.ptr, B.ptr, self.plan.ptr, stream))

    def sub(self, A, B, stream=None):
        return _check(lib().fa_op_sub(self.ptr, A.ptr, B.ptr, self.plan.ptr, stream))

    def neg(self, A, stream=None):
        return _check(lib().fa_op_neg(self.ptr, A.ptr, self.plan.ptr, stream))

    def ifft(self, stream=None):
        return _check(lib().fa_ifft(self.ptr, self.plan.ptr, stream))

    def export_mpn(self, z_ptr, zn, stream=None):
        return _check(lib().fa_export_mpn(z_ptr, zn, self.ptr, self.plan.ptr, stream))

    def export_mpn_signed(self, z_ptr, zn, sign_ptr, nslots, stream=None):
        return _check(lib().fa_export_mpn_signed(z_ptr, zn, sign_ptr, self.ptr, nslots,
                                                 self.plan.ptr, stream))

    def export_mpn_signed_trunc(self, z_ptr, zn, sign_ptr, nslots, lo, stream=None):
        """limbs [lo, lo + zn) of |value| with the reference's truncation"""
        return _check(lib().fa_export_mpn_signed_trunc(z_ptr, zn, sign_ptr, self.ptr, nslots,
                                                       lo, self.plan.ptr, stream))

    def export_nmod(self, z_ptr, n, zl=None, zh=None, stream=None):
        if zl is None:
            return _check(lib().fa_export_nmod(z_ptr, self.ptr, nmod_init(n), self.plan.ptr,
                                               stream))
        return _check(lib().fa_export_nmod_range(z_ptr, self.ptr, zl, zh, nmod_init(n),
                                                 self.plan.ptr, stream))


def nmod_init(n):
    """nmod_init (nmod.h:284): n, ninv = floor((2^128-1)/(n<<norm)) - 2^64, norm."""
    n = int(n)
    norm = 64 - n.bit_length()
    ninv = ((1 << 128) - 1) // (n << norm) - (1 << 64)
    return NmodT(n, ninv & ((1 << 64) - 1), norm)


def nmod_poly_mul_mid(a, b, zl, zh, n):
    """_nmod_poly_mul_mid_default_mpn_ctx: coefficients [zl, zh) of a*b mod n."""
    a = _arr(a)
    b = _arr(b)
    z = np.zeros(max(zh - zl, 1), dtype=np.uint64)
    _check(lib().fa_nmod_poly_mul_mid(_p(z), zl, zh, _p(a), len(a), _p(b), len(b), nmod_init(n)))
    return z[: max(zh - zl, 0)]


def nmod_poly_mul_mid_device(z_ptr, zl, zh, a_ptr, an, b_ptr, bn, n, stream=None):
    """_nmod_poly_mul_mid on device arrays, enqueued on `stream`"""
    return _check(lib().fa_nmod_poly_mul_mid_device(z_ptr, zl, zh, a_ptr, an, b_ptr, bn,
                                                    nmod_init(n), stream))


def nmod_poly_mulmid_repack(a, b, znlo, zn, n):
    """the small-modulus packed route: (1, z) when it applied, (0, None) if not"""
    a = _arr(a)
    b = a if b is a else _arr(b)
    z = np.zeros(max(zn - znlo, 1), dtype=np.uint64)
    rc = _check(lib().fa_nmod_poly_mulmid_repack(_p(z), _p(a), len(a), _p(b), len(b), znlo, zn,
                                                 nmod_init(n)))
    return (1, z[: zn - znlo]) if rc == 1 else (0, None)


def nmod_poly_mulmid_fft_small(a, b, znlo, zn, n):
    """_nmod_poly_mulmid_fft_small: packed route for n <= 23 when it applies"""
    a = _arr(a)
    b = a if b is a else _arr(b)
    z = np.zeros(max(zn - znlo, 1), dtype=np.uint64)
    _check(lib().fa_nmod_poly_mulmid_fft_small(_p(z), _p(a), len(a), _p(b), len(b), znlo, zn,
                                               nmod_init(n)))
    return z[: zn - znlo]


def nmod_poly_inv_series(f, m, n):
    """f^-1 mod x^m over Z/nZ (f[0] invertible)"""
    f = _arr(f)
    g = np.empty(m, dtype=np.uint64)
    _check(lib().fa_nmod_poly_inv_series(_p(g), _p(f), len(f), m, nmod_init(n)))
    return g


def nmod_poly_divrem(a, b, n):
    """_nmod_poly_divrem_mpn_ctx: (q, r) with a = b q + r, len(r) = len(b) - 1"""
    a = _arr(a)
    b = _arr(b)
    q = np.empty(len(a) - len(b) + 1, dtype=np.uint64)
    r = np.empty(max(len(b) - 1, 1), dtype=np.uint64)
    _check(lib().fa_nmod_poly_divrem(_p(q), _p(r), _p(a), len(a), _p(b), len(b), nmod_init(n)))
    return q, r[: len(b) - 1]


class DivremPrecomp:
    """nmod_poly_divrem_precomp: divisor b and its inverse to precision Bn
    kept on the device"""

    def __init__(self, b, Bn, n):
        b = _arr(b)
        self.h = _VP()
        self.n = n
        self.bn = len(b)
        _check(lib().fa_nmod_poly_divrem_precomp_init(ctypes.byref(self.h), _p(b), len(b), Bn,
                                                      nmod_init(n)))

    def divrem(self, a):
        """(q, r), or None when len(a) - len(b) + 1 exceeds Bn"""
        a = _arr(a)
        q = np.empty(len(a) - self.bn + 1, dtype=np.uint64)
        r = np.empty(max(self.bn - 1, 1), dtype=np.uint64)
        rc = _check(lib().fa_nmod_poly_divrem_precomp(_p(q), _p(r), _p(a), len(a), self.h,
                                                      nmod_init(self.n)))
        return (q, r[: self.bn - 1]) if rc == 1 else None

    def close(self):
        if self.h:
            lib().fa_nmod_poly_divrem_precomp_clear(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MulPrecomp:
    """mul_precomp_struct: b transformed once on a cyclic plan of 2^depth
    points, for middle products a b on windows the transform holds"""

    def __init__(self, b, depth, n, btrunc=0):
        b = _arr(b)
        self.h = _VP()
        self.n = n
        _check(lib().fa_nmod_poly_mul_precomp_init(ctypes.byref(self.h), _p(b), len(b), btrunc,
                                                   depth, nmod_init(n)))

    def mul_mid(self, a, zl, zh):
        """coefficients [zl, zh) of a b, or None when the window does not fit"""
        a = _arr(a)
        z = np.empty(max(zh - zl, 1), dtype=np.uint64)
        rc = _check(lib().fa_nmod_poly_mul_mid_precomp(_p(z), zl, zh, _p(a), len(a), self.h,
                                                       nmod_init(self.n)))
        return z[: zh - zl] if rc == 1 else None

    def close(self):
        if self.h:
            lib().fa_nmod_poly_mul_precomp_clear(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fmpz_poly_mul_mid(a, aw, b, bw, zl, zh, zw):
    """_fmpz_poly_mul_mid_default_mpn_ctx on w-limb two's complement
    coefficients.  Returns the (zh-zl)*zw output limbs, or None where the
    reference declines (more than 8 primes, fmpz_poly_mul.c:377-379)."""
    a = _arr(a)
    b = _arr(b)
    z = np.zeros(max((zh - zl) * zw, 1), dtype=np.uint64)
    rc = _check(lib().fa_fmpz_poly_mul_mid(_p(z), zw, zl, zh, _p(a), aw, len(a) // aw,
                                           _p(b), bw, len(b) // bw))
    return z[: (zh - zl) * zw] if rc == 1 else None


def fmpz_poly_mul_ks(a, aw, b, bw, zl, zh, zw):
    """fmpz product by Kronecker substitution onto the integer product: the
    route taken when fmpz_poly_mul_mid declines; always computes"""
    a = _arr(a)
    b = _arr(b)
    z = np.zeros(max((zh - zl) * zw, 1), dtype=np.uint64)
    _check(lib().fa_fmpz_poly_mul_ks(_p(z), zw, zl, zh, _p(a), aw, len(a) // aw, _p(b), bw,
                                     len(b) // bw))
    return z[: (zh - zl) * zw]


def fmpz_vec_max_bits(coeffs, w):
    """_fmpz_vec_max_bits of w-limb two's complement coefficients"""
    c = _arr(coeffs)
    return int(lib().fa_fmpz_vec_max_bits(c.ctypes.data, w, len(c) // w))


def batch_mul(requests, ks=False):
    """Runs a list of host-side products as one batch (fa_batch_mul).

    Each request is a dict: {"kind": "mpn", "a": limbs, "b": limbs},
    {"kind": "nmod", "a", "b", "n", "zl", "zh"} or
    {"kind": "fmpz", "a", "aw", "b", "bw", "zl", "zh", "zw"} (coefficients as
    flat w-limb two's complement arrays).  Returns a list of (status, output)
    with output None unless status == 1.  ks=True computes the fmpz products
    the eight-prime rule declines by Kronecker substitution."""
    n = len(requests)
    R = (Request * max(n, 1))()
    keep = []
    outs = []
    for k, q in enumerate(requests):
        r = R[k]
        a = _arr(q["a"])
        b = _arr(q["b"])
        keep += [a, b]
        kind = q["kind"]
        if kind == "mpn":
            r.kind = REQ_MPN
            r.an, r.bn = len(a), len(b)
            z = np.zeros(len(a) + len(b), dtype=np.uint64)
            r.np = q.get("np", 0)
            r.bits = q.get("bits", 0)
            r.trunc = q.get("trunc", TRUNC_AUTO)
        elif kind == "nmod":
            r.kind = REQ_NMOD
            r.an, r.bn = len(a), len(b)
            r.zl, r.zh = q["zl"], q["zh"]
            r.mod = nmod_init(q["n"])
            z = np.zeros(max(q["zh"] - q["zl"], 1), dtype=np.uint64)
        elif kind == "fmpz":
            r.kind = REQ_FMPZ
            r.aw, r.bw, r.zw = q["aw"], q["bw"], q["zw"]
            r.an, r.bn = len(a) // q["aw"], len(b) // q["bw"]
            r.zl, r.zh = q["zl"], q["zh"]
            r.abits, r.bbits = q.get("abits", 0), q.get("bbits", 0)
            z = np.zeros(max((q["zh"] - q["zl"]) * q["zw"], 1), dtype=np.uint64)
        else:
            raise ValueError("kind must be mpn, nmod or fmpz")
        r.a = a.ctypes.data
        r.b = b.ctypes.data
        r.z = z.ctypes.data
        outs.append(z)
    _check(lib().fa_batch_mul(ctypes.cast(R, _VP), n, BATCH_FMPZ_KS if ks else 0))
    res = []
    for k, q in enumerate(requests):
        st = R[k].status
        z = outs[k]
        if q["kind"] == "nmod":
            z = z[: max(q["zh"] - q["zl"], 0)]
        elif q["kind"] == "fmpz":
            z = z[: (q["zh"] - q["zl"]) * q["zw"]]
        res.append((st, z if st == 1 else None))
    return res


class DeviceBatch:
    """A batch over device buffers, planned once and replayed
    (fa_batch_create with FA_BATCH_DEVICE, optionally FA_BATCH_GRAPH).
    `reqs` is a ctypes array of Request whose a, b, z are device addresses;
    it must stay alive as long as the batch."""

    def __init__(self, reqs, graph=True, ks=False):
        self.reqs = reqs
        self.h = _VP()
        flags = BATCH_DEVICE | (BATCH_GRAPH if graph else 0) | (BATCH_FMPZ_KS if ks else 0)
        _check(lib().fa_batch_create(ctypes.byref(self.h), ctypes.cast(reqs, _VP), len(reqs),
                                     flags))

    def launch(self, stream=None):
        _check(lib().fa_batch_launch(self.h, stream))

    def close(self):
        if self.h:
            lib().fa_batch_destroy(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_headers(count, seed=0xF1A70005):
    """the first `count` headers of the device-generated mixed stream"""
    H = (StreamReq * count)()
    _check(lib().fa_stream_headers(seed, count, ctypes.cast(H, _VP)))
    return H


def stream_word(key, j):
    """word j of the SplitMix64 sequence seeded with key"""
    return int(lib().fa_stream_word(key, j))


def stream_fill_device(specs, stream=None):
    """writes the operands of a FillSpec array on the device and waits"""
    _check(lib().fa_stream_fill_device(ctypes.cast(specs, _VP), len(specs), stream))


def eval_device(specs, stream=None):
    """values of an EvalSpec array (numpy uint64)"""
    out = np.zeros(max(len(specs), 1), dtype=np.uint64)
    _check(lib().fa_eval_device(ctypes.cast(specs, _VP), len(specs), out.ctypes.data, stream))
    return out[:len(specs)]


def mpn_candidates(an, bn):
    """the 64 candidate plans (np, bits, trunc) of an x bn limbs; (0, 0, 0) = invalid"""
    out = np.zeros((64, 3), dtype=np.int32)
    _check(lib().fa_mpn_candidates(an, bn, out.ctypes.data))
    return out


def plan_features(an, bn):
    f = np.zeros(16, dtype=np.float32)
    _check(lib().fa_plan_features(an, bn, f.ctypes.data))
    return f


def autotune_mpn(an, bn, reps=3):
    """device time (ms) of every candidate plan; -1 where invalid"""
    ms = np.zeros(64, dtype=np.float32)
    _check(lib().fa_autotune_mpn(an, bn, reps, ms.ctypes.data))
    return ms


class MLP:
    """The planner network (fa_mlp_*): 16 -> 64 -> 64 -> 64, FP32 on the GPU."""

    def __init__(self, seed=1):
        self.h = _VP()
        _check(lib().fa_mlp_create(ctypes.byref(self.h), seed))
        self.nparam = lib().fa_mlp_param_count()

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        _check(lib().fa_mlp_get_params(self.h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().fa_mlp_set_params(self.h, p.ctypes.data))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), 64), dtype=np.float32)
        _check(lib().fa_mlp_forward(self.h, X.ctypes.data, len(X), Y.ctypes.data))
        return Y

    def loss_grad(self, X, T, M):
        X, T, M = (np.ascontiguousarray(v, dtype=np.float32) for v in (X, T, M))
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        _check(lib().fa_mlp_loss_grad(self.h, X.ctypes.data, T.ctypes.data, M.ctypes.data,
                                      len(X), loss.ctypes.data, g.ctypes.data))
        return float(loss[0]), g

    def train(self, X, T, M, batch=256, epochs=10, lr=1e-3):
        X, T, M = (np.ascontiguousarray(v, dtype=np.float32) for v in (X, T, M))
        losses = np.zeros(epochs, dtype=np.float32)
        _check(lib().fa_mlp_train(self.h, X.ctypes.data, T.ctypes.data, M.ctypes.data, len(X),
                                  batch, epochs, ctypes.c_float(lr), losses.ctypes.data))
        return losses

    def plan_mpn(self, shapes):
        """learned plans: n x (np, bits, trunc)"""
        shapes = np.ascontiguousarray(shapes, dtype=np.int64)
        out = np.zeros((len(shapes), 3), dtype=np.int32)
        _check(lib().fa_mlp_plan_mpn(self.h, shapes.ctypes.data, len(shapes), out.ctypes.data))
        return out

    def close(self):
        if self.h:
            lib().fa_mlp_destroy(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SdFft:
    """sd_fft_ctx_t with sd_fft_trunc / sd_ifft_trunc (fft_small.h:206-281)."""

    def __init__(self, p):
        self.ctx = SdFftCtx()
        _check(lib().fa_sd_fft_ctx_init_prime(ctypes.byref(self.ctx), p))

    def __del__(self):
        try:
            lib().fa_sd_fft_ctx_clear(ctypes.byref(self.ctx))
        except Exception:
            pass

    def w(self, j):
        return lib().fa_sd_fft_ctx_w(ctypes.byref(self.ctx), j)

    def fft_trunc(self, d, L, itrunc, otrunc):
        d = np.ascontiguousarray(d, dtype=np.float64).copy()
        _check(lib().fa_sd_fft_trunc(ctypes.byref(self.ctx), d.ctypes.data_as(_PD), L, itrunc,
                                     otrunc))
        return d

    def ifft_trunc(self, d, L, trunc):
        d = np.ascontiguousarray(d, dtype=np.float64).copy()
        _check(lib().fa_sd_ifft_trunc(ctypes.byref(self.ctx), d.ctypes.data_as(_PD), L, trunc))
        return d

    def point_mul(self, a, b, m, depth):
        """sd_fft_ctx_point_mul: a m b pointwise mod p, in [0, p)"""
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().fa_sd_fft_ctx_point_mul(ctypes.byref(self.ctx), a.ctypes.data_as(_PD),
                                             b.ctypes.data_as(_PD), m, depth))
        return a

    def point_sqr(self, a, m, depth):
        """sd_fft_ctx_point_sqr: a m a pointwise mod p, in [0, p)"""
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        _check(lib().fa_sd_fft_ctx_point_sqr(ctypes.byref(self.ctx), a.ctypes.data_as(_PD), m,
                                             depth))
        return a
