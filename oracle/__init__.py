"""ctypes bindings of the CPU oracle (test infrastructure only).

The oracle is a C restatement of the reference denoiser
(src/core/nldct.cpp) and TV-L1 flow (lib/tvl1flow), built by
oracle/Makefile into liboracle.so.  It is the checker for the GPU path: only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.

`ref_tvl1()` loads oracle/_ref/libtvl1ref.so, the reference's own
lib/tvl1flow/tvl1flow_lib.c compiled from its sources, when it was built.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    """Field-for-field mirror of bm4dof_params (include/bm4dof.h), itself a
    mirror of struct nlbParams (src/core/nldct.h:55-82)."""
    _fields_ = [
        ("sigma", C.c_float),
        ("sizePatch", C.c_uint32),
        ("sizePatchTime", C.c_uint32),
        ("nSimilarPatches", C.c_uint32),
        ("sizeSearchWindow", C.c_uint32),
        ("sizeSearchTimeRangeFwd", C.c_uint32),
        ("sizeSearchTimeRangeBwd", C.c_uint32),
        ("boundary", C.c_uint32),
        ("offSet", C.c_uint32),
        ("offSetTime", C.c_uint32),
        ("beta", C.c_float),
        ("betaMean", C.c_float),
        ("tau", C.c_float),
        ("orderInvariance", C.c_int32),
        ("isFirstStep", C.c_int32),
        ("doPasteBoost", C.c_int32),
        ("verbose", C.c_int32),
        ("colorSpace", C.c_int32),
        ("transform", C.c_int32),
        ("sizeSearchWindowPred", C.c_uint32),
        ("nSimilarPatchesPred", C.c_uint32),
        ("dsub", C.c_float),
        ("agg_window", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER(Params)
        L.or_init_params.argtypes = [P, C.c_uint, C.c_float, C.c_uint, C.c_int]
        L.or_set_search_window.argtypes = [P, C.c_uint]
        L.or_set_patch_size.argtypes = [P, C.c_uint, C.c_uint]
        L.or_set_n_similar.argtypes = [P, C.c_uint]
        L.or_last_error.restype = C.c_char_p
        L.or_color_transform.argtypes = [_f32p, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_int]
        L.or_psnr.argtypes = [_f32p, _f32p, C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_determine_factor.argtypes = [C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
        L.or_tiles.argtypes = [C.c_uint, C.c_uint, C.c_int, C.c_int, _i32p]
        L.or_partial_sort.argtypes = [_f32p, _u32p, C.c_long, C.c_long, C.c_int]
        L.or_bases.argtypes = [P, _f32p, _f32p, _f32p]
        L.or_set_predictive.argtypes = [C.c_int]
        L.or_filter_group.argtypes = [P, C.c_int, C.c_int, _f32p, C.c_void_p, C.POINTER(C.c_float)]
        L.or_run.argtypes = [P, P, _f32p, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                             C.c_void_p, C.c_void_p, _f32p, C.c_void_p, C.c_int, C.c_int,
                             C.POINTER(C.c_float)]
        L.or_bm.restype = C.c_long
        L.or_bm.argtypes = [P, _f32p, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_void_p,
                            C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_void_p]
        # TV-L1
        L.or_tv_flow.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float,
                                 C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.c_int,
                                 C.c_float, _i32p]
        L.or_tv_solve.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float,
                                  C.c_float, C.c_float, C.c_int, C.c_float, _i32p]
        L.or_tv_gaussian.argtypes = [_f32p, C.c_int, C.c_int, C.c_double]
        L.or_tv_gauss_taps.argtypes = [C.c_double, _f64p, C.c_int]
        L.or_tv_zoom_out.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_float]
        L.or_tv_zoom_in.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_tv_zoom_size.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.c_float]
        L.or_tv_bicubic_at.restype = C.c_float
        L.or_tv_bicubic_at.argtypes = [_f32p, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int]
        for name in ("or_tv_divergence",):
            getattr(L, name).argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int]
        for name in ("or_tv_forward_gradient", "or_tv_centered_gradient"):
            getattr(L, name).argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int]
        _LIB = L
    return _LIB


def ref_tvl1():
    """The reference's own TV-L1 library (oracle/_ref), or None when absent."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libtvl1ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.Dual_TVL1_optic_flow_multiscale.restype = None
        L.Dual_TVL1_optic_flow_multiscale.argtypes = [
            _f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
            C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_bool]
        L.gaussian.argtypes = [_f32p, C.c_int, C.c_int, C.c_double]
        L.zoom_out.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_float]
        L.zoom_in.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.bicubic_interpolation_at.restype = C.c_float
        L.bicubic_interpolation_at.argtypes = [_f32p, C.c_float, C.c_float, C.c_int, C.c_int,
                                               C.c_bool]
        L.Dual_TVL1_optic_flow.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int,
                                           C.c_float, C.c_float, C.c_float, C.c_int, C.c_float,
                                           C.c_bool]
        _REF = L
    return _REF


# ---------------------------------------------------------------- parameters

def init_params(step, sigma, frames, verbose=0):
    p = Params()
    lib().or_init_params(C.byref(p), step, sigma, frames, verbose)
    return p


def cli_params(step, sigma, frames, px=-1, pt=1, wx=-1, np_=-1, beta=-1.0, beta_mean=-1.0,
               tau=-1.0, order_inv=False, no_step=False):
    """Parameters as the bm4d-of command line builds them
    (src/main_vnlb.cpp:206-246).  `tau` is only applied under VBM3D_SEARCH in
    the reference (main_vnlb.cpp:227-236), so it is ignored here too."""
    L = lib()
    p = init_params(step, sigma, frames)
    if wx >= 0:
        L.or_set_search_window(C.byref(p), wx)
    if px >= 0:
        L.or_set_patch_size(C.byref(p), frames, px)
    if np_ >= 0:
        L.or_set_n_similar(C.byref(p), np_)
    if beta >= 0:
        p.beta = beta
        p.betaMean = beta
    if beta_mean >= 0:
        p.betaMean = beta_mean
    if pt >= 0:
        p.sizePatchTime = pt
    if no_step:
        p.offSet = 1
    if order_inv:
        p.orderInvariance = 1
    return p


# ------------------------------------------------------------------ denoiser

def run(p1, p2, noisy, fflow=None, bflow=None, basic=None, nparts=2, tie_mode=0):
    """runNlBayes on a (F, C, H, W) float32 array.  Returns (basic, final,
    groups_ratio); final is None when step 2 is skipped."""
    L = lib()
    noisy = np.ascontiguousarray(noisy, np.float32)
    F, Cc, H, W = noisy.shape
    b = np.zeros_like(noisy) if basic is None else np.ascontiguousarray(basic, np.float32).copy()
    fin = np.zeros_like(noisy)
    gr = (C.c_float * 2)()
    ff = bf = None
    if fflow is not None:
        ff = np.ascontiguousarray(fflow, np.float32)
        bf = np.ascontiguousarray(bflow, np.float32)
    rc = L.or_run(C.byref(p1), C.byref(p2), noisy, W, H, F, Cc,
                  ff.ctypes.data if ff is not None else None,
                  bf.ctypes.data if bf is not None else None, b,
                  fin.ctypes.data if p2.sizePatch else None, nparts, tie_mode, gr)
    if rc:
        raise RuntimeError(L.or_last_error().decode())
    return b, (fin if p2.sizePatch else None), (gr[0], gr[1])


def bm(p, image, fflow=None, bflow=None, nparts=2, tie_mode=0):
    """Neighbour lists at every grid position (see or_bm)."""
    L = lib()
    image = np.ascontiguousarray(image, np.float32)
    F, Cc, H, W = image.shape
    ff = bf = None
    if fflow is not None:
        ff = np.ascontiguousarray(fflow, np.float32)
        bf = np.ascontiguousarray(bflow, np.float32)
    fa = ff.ctypes.data if ff is not None else None
    ba = bf.ctypes.data if bf is not None else None
    n = L.or_bm(C.byref(p), image, W, H, F, Cc, fa, ba, nparts, tie_mode, None, None, None, None)
    if n < 0:
        raise RuntimeError(L.or_last_error().decode())
    npm = p.nSimilarPatches
    idx = np.zeros(n * npm, np.uint32)
    nsim = np.zeros(n, np.uint32)
    ref = np.zeros(n, np.uint32)
    tile = np.zeros(n, np.uint32)
    L.or_bm(C.byref(p), image, W, H, F, Cc, fa, ba, nparts, tie_mode, idx.ctypes.data,
            nsim.ctypes.data, ref.ctypes.data, tile.ctypes.data)
    return idx.reshape(n, npm), nsim, ref, tile


def set_predictive(on):
    """The reference's VBM3D_SEARCH predictive search in block matching
    (nldct.cpp:904-1130, 1368-1589) on (1) or off (0, the reference's
    default build)."""
    lib().or_set_predictive(1 if on else 0)


def bases(p):
    P = p.sizePatch * p.sizePatch * p.sizePatchTime
    B = np.zeros(P * P, np.float32)
    Bi = np.zeros(P * P, np.float32)
    w = np.zeros(P, np.float32)
    if lib().or_bases(C.byref(p), B, Bi, w) < 0:
        raise RuntimeError(lib().or_last_error().decode())
    return B.reshape(P, P), Bi.reshape(P, P), w


def filter_group(p, gn, gb=None):
    """computeBayesEstimateStep{1,2} on one group of shape (C, P, N) ->
    (filtered group, weight, early_out)."""
    g = np.ascontiguousarray(gn, np.float32).copy()
    Cc, P, N = g.shape
    w = C.c_float(0)
    gbp = None
    if gb is not None:
        gb = np.ascontiguousarray(gb, np.float32)
        gbp = gb.ctypes.data
    r = lib().or_filter_group(C.byref(p), Cc, N, g.reshape(-1), gbp, C.byref(w))
    if r < 0:
        raise RuntimeError(lib().or_last_error().decode())
    return g, w.value, bool(r)


def tiles(W, H, boundary, nparts):
    out = np.zeros(8 * nparts, np.int32)
    if lib().or_tiles(W, H, boundary, nparts, out):
        raise RuntimeError(lib().or_last_error().decode())
    return out.reshape(nparts, 8)


def psnr(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    b = np.ascontiguousarray(b, np.float32).reshape(-1)
    p, r = C.c_float(), C.c_float()
    lib().or_psnr(a, b, a.size, C.byref(p), C.byref(r))
    return p.value, r.value


def color_transform(video, forward):
    v = np.ascontiguousarray(video, np.float32).copy()
    F, Cc, H, W = v.shape
    lib().or_color_transform(v.reshape(-1), W, H, F, Cc, 1 if forward else 0)
    return v


def partial_sort(d, idx, middle, tie_mode=0):
    d = np.ascontiguousarray(d, np.float32).copy()
    idx = np.ascontiguousarray(idx, np.uint32).copy()
    lib().or_partial_sort(d, idx, middle, d.size, tie_mode)
    return d, idx


# --------------------------------------------------------------------- TV-L1

def tvl1(I0, I1, tau=0.25, lam=0.15, theta=0.3, nscales=5, fscale=0, zfactor=0.5, warps=10,
         eps=0.01):
    """Oracle TV-L1 on two (H, W) frames -> (u, v, iters[nscales, warps])."""
    I0 = np.ascontiguousarray(I0, np.float32)
    I1 = np.ascontiguousarray(I1, np.float32)
    H, W = I0.shape
    u = np.zeros((H, W), np.float32)
    v = np.zeros((H, W), np.float32)
    it = np.zeros(nscales * warps, np.int32)
    lib().or_tv_flow(I0, I1, u, v, W, H, tau, lam, theta, nscales, fscale, zfactor, warps, eps,
                     it)
    return u, v, it.reshape(nscales, warps)


def _omp_threads(n):
    """Set the OpenMP thread count of this process; returns the previous one."""
    gomp = C.CDLL("libgomp.so.1")
    prev = gomp.omp_get_max_threads()
    gomp.omp_set_num_threads(int(n))
    return prev


def tvl1_reference(I0, I1, tau=0.25, lam=0.15, theta=0.3, nscales=5, fscale=0, zfactor=0.5,
                   warps=10, eps=0.01):
    """The reference library's own TV-L1 (oracle/_ref); None when absent.
    Run with one OpenMP thread for a deterministic error reduction."""
    R = ref_tvl1()
    if R is None:
        return None
    I0 = np.ascontiguousarray(I0, np.float32).copy()
    I1 = np.ascontiguousarray(I1, np.float32).copy()
    H, W = I0.shape
    u = np.zeros((H, W), np.float32)
    v = np.zeros((H, W), np.float32)
    prev = _omp_threads(1)
    try:
        R.Dual_TVL1_optic_flow_multiscale(I0, I1, u, v, W, H, tau, lam, theta, nscales, fscale,
                                          zfactor, warps, eps, False)
    finally:
        _omp_threads(prev)
    return u, v
