"""Python binding of the MI355X BM4D-OF denoiser's C-ABI (include/bm4dof.h).

The compute path is libbm4dof.so (HIP kernels for gfx950, built by
`make -C bm4d-of_amd`).  There is no CPU fallback: loading fails loudly when
the library is missing, and every call reports the library's error text.

Videos are numpy arrays of shape (F, C, H, W), float32, the reference's
planar layout (src/utils/lib_videot.hpp:88-92); flows are (F, 2, H, W).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbm4dof.so")
IO_PATH = os.path.join(_HERE, "libbm4dof_io.so")
VIDEONLB_PATH = os.path.join(_HERE, "libbm4dof_videonlb.so")
# command-line tools: bm4d-of (src/main_vnlb.cpp), tvl1flow
# (lib/tvl1flow/main.c) and tvl1flow-seq (scripts/tvl1flow-seq.sh)
TOOLS = {name: os.path.join(_HERE, "bin", name) for name in ("bm4d-of", "tvl1flow", "tvl1flow-seq")}
IO_EXPORTS = ["bm4dof_io_error", "bm4dof_io_free", "bm4dof_io_read", "bm4dof_io_read_gray",
              "bm4dof_io_write"]
_LIB = None
_IO = None

BM4DOF_RGB, BM4DOF_YUV = 0, 1
BM4DOF_DCT, BM4DOF_BIOR1_5 = 0, 1


class Params(C.Structure):
    """bm4dof_params: fixed-width mirror of struct nlbParams
    (src/core/nldct.h:55-82)."""
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


class TVL1Params(C.Structure):
    """bm4dof_tvl1_params, in tvl1flow's argument order (lib/tvl1flow/main.c:86-99)."""
    _fields_ = [("tau", C.c_float), ("lambda_", C.c_float), ("theta", C.c_float),
                ("nscales", C.c_int), ("fscale", C.c_int), ("zfactor", C.c_float),
                ("warps", C.c_int), ("epsilon", C.c_float), ("verbose", C.c_int)]


# Every symbol include/bm4dof.h declares.
EXPORTS = [
    "bm4dof_init_params", "bm4dof_set_search_window", "bm4dof_set_patch_size",
    "bm4dof_set_n_similar", "bm4dof_ctx_create", "bm4dof_ctx_destroy", "bm4dof_last_error",
    "bm4dof_global_error", "bm4dof_run", "bm4dof_run_dev", "bm4dof_bm_count", "bm4dof_bm",
    "bm4dof_last_timings", "bm4dof_last_stats", "Dual_TVL1_optic_flow_multiscale", "bm4dof_tvl1",
    "bm4dof_tvl1_sequence", "bm4dof_ctx_set_option", "bm4dof_mlp_set", "bm4dof_mlp_get",
    "bm4dof_mlp_forward", "bm4dof_mlp_grad", "bm4dof_mlp_train", "bm4dof_learned_pairs",
    "bm4dof_run_learned", "bm4dof_run_learned_dev", "bm4dof_batch_create", "bm4dof_batch_run",
    "bm4dof_batch_last_ms", "bm4dof_batch_report", "bm4dof_batch_stats", "bm4dof_batch_destroy",
    "bm4dof_autotune", "bm4dof_tvl1_noalloc",
]

BATCH_GRAPH = 1
BATCH_AUTOTUNE = 2

MLP_NPARAMS = 33344


class Seq(C.Structure):
    """bm4dof_seq: one gray sequence of a batch (device pointers)."""
    _fields_ = [("noisy", C.c_void_p), ("fflow", C.c_void_p), ("bflow", C.c_void_p),
                ("basic", C.c_void_p), ("final", C.c_void_p), ("w", C.c_uint), ("h", C.c_uint),
                ("f", C.c_uint)]


class BM4DOFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bm4dof error {code}: {msg}")
        self.code = code


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    """Load libbm4dof.so (raises when it is missing: no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}`")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7, loaded under another file name).  Loading torch first
    # lets the dynamic loader bind this library's libamdhip64.so.7 to that
    # same runtime; the other order would load a second runtime, and torch's
    # would then find no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(Params)
    vp = C.c_void_p
    L.bm4dof_init_params.argtypes = [P, C.c_uint, C.c_float, C.c_uint, C.c_uint, C.c_uint,
                                     C.c_uint, C.c_int]
    L.bm4dof_init_params.restype = None
    L.bm4dof_set_search_window.argtypes = [P, C.c_uint]
    L.bm4dof_set_patch_size.argtypes = [P, C.c_uint, C.c_uint]
    L.bm4dof_set_n_similar.argtypes = [P, C.c_uint]
    for f in ("bm4dof_set_search_window", "bm4dof_set_patch_size", "bm4dof_set_n_similar"):
        getattr(L, f).restype = None
    L.bm4dof_ctx_create.argtypes = [C.POINTER(vp), C.c_int]
    L.bm4dof_ctx_destroy.argtypes = [vp]
    L.bm4dof_ctx_destroy.restype = None
    L.bm4dof_last_error.argtypes = [vp]
    L.bm4dof_last_error.restype = C.c_char_p
    L.bm4dof_global_error.restype = C.c_char_p
    run_args = [vp, P, P, vp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp, vp, vp, vp, C.c_int,
                C.POINTER(C.c_float)]
    L.bm4dof_run.argtypes = run_args
    L.bm4dof_run_dev.argtypes = run_args
    L.bm4dof_bm_count.argtypes = [vp, P, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_int,
                                  C.POINTER(C.c_int64)]
    L.bm4dof_bm.argtypes = [vp, P, vp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp, vp, C.c_int,
                            vp, vp, vp, vp]
    L.bm4dof_last_timings.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_char_p), C.c_int]
    L.bm4dof_last_stats.argtypes = [vp, C.POINTER(C.c_int64), C.c_int]
    L.bm4dof_ctx_set_option.argtypes = [vp, C.c_char_p, C.c_longlong]
    L.bm4dof_mlp_set.argtypes = [vp, vp]
    L.bm4dof_mlp_get.argtypes = [vp, vp]
    L.bm4dof_mlp_forward.argtypes = [vp, vp, C.c_int64, vp, C.c_int]
    L.bm4dof_mlp_grad.argtypes = [vp, vp, vp, C.c_int, C.POINTER(C.c_float), vp]
    L.bm4dof_mlp_train.argtypes = [vp, vp, vp, C.c_int64, C.c_int, C.c_int, C.c_float, C.c_uint, vp]
    L.bm4dof_learned_pairs.argtypes = [vp, P, vp, vp, C.c_uint, C.c_uint, C.c_uint, vp, vp, C.c_int,
                                       C.c_int, C.c_float, vp, vp, vp, C.c_int64,
                                       C.POINTER(C.c_int64)]
    L.bm4dof_batch_create.argtypes = [vp, C.POINTER(Seq), P, P, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_void_p)]
    L.bm4dof_batch_run.argtypes = [vp]
    L.bm4dof_batch_last_ms.argtypes = [vp]
    L.bm4dof_batch_last_ms.restype = C.c_float
    L.bm4dof_batch_report.argtypes = [vp]
    L.bm4dof_batch_report.restype = C.c_char_p
    L.bm4dof_batch_stats.argtypes = [vp, C.POINTER(C.c_int64), C.c_int]
    L.bm4dof_batch_destroy.argtypes = [vp]
    L.bm4dof_batch_destroy.restype = None
    L.bm4dof_autotune.argtypes = [vp, P, P, vp, C.c_uint, C.c_uint, C.c_uint, vp, vp, vp, vp, C.c_int,
                                  C.c_char_p, C.c_int]
    learned = [vp, P, P, vp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp, vp, vp, C.c_int, C.c_float,
               C.POINTER(C.c_float)]
    L.bm4dof_run_learned.argtypes = learned
    L.bm4dof_run_learned_dev.argtypes = learned
    if not hasattr(L, "bm4dof_tvl1"):
        _LIB = L
        return L
    L.Dual_TVL1_optic_flow_multiscale.restype = None
    L.Dual_TVL1_optic_flow_multiscale.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_float,
                                                  C.c_float, C.c_float, C.c_int, C.c_int,
                                                  C.c_float, C.c_int, C.c_float, C.c_bool]
    L.bm4dof_tvl1.argtypes = [vp, C.POINTER(TVL1Params), vp, vp, vp, vp, C.c_int, C.c_int,
                              C.c_int, vp]
    L.bm4dof_tvl1_noalloc.argtypes = [vp, C.POINTER(TVL1Params), vp, vp, vp, vp, C.c_int, C.c_int, vp]
    L.bm4dof_tvl1_sequence.argtypes = [vp, C.POINTER(TVL1Params), vp, C.c_int, C.c_int,
                                       C.c_int, vp, vp, C.c_int]
    _LIB = L
    return L


# ---------------------------------------------------------------- parameters

def init_params(step, sigma, width, height, frames, channels=1, verbose=0):
    p = Params()
    lib().bm4dof_init_params(C.byref(p), step, sigma, width, height, frames, channels, verbose)
    return p


def cli_params(step, sigma, shape, px=-1, pt=1, wx=-1, np_=-1, beta=-1.0, beta_mean=-1.0,
               order_inv=False, no_step=False):
    """Parameters as the bm4d-of command line builds them from its flags
    (src/main_vnlb.cpp:206-246).  shape = (F, C, H, W)."""
    F, Cc, H, W = shape
    L = lib()
    p = init_params(step, sigma, W, H, F, Cc)
    if wx >= 0:
        L.bm4dof_set_search_window(C.byref(p), wx)
    if px >= 0:
        L.bm4dof_set_patch_size(C.byref(p), F, px)
    if np_ >= 0:
        L.bm4dof_set_n_similar(C.byref(p), np_)
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


# ------------------------------------------------------------------ context

class Context:
    """A bm4dof_ctx bound to one HIP device."""

    def __init__(self, device=0):
        self.device = device
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.bm4dof_ctx_create(C.byref(h), device)
        if rc:
            raise BM4DOFError(rc, self._L.bm4dof_global_error().decode())
        self.h = h

    def close(self):
        if self.h:
            self._L.bm4dof_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise BM4DOFError(rc, self._L.bm4dof_last_error(self.h).decode())

    def run(self, p1, p2, noisy, fflow=None, bflow=None, basic=None, nparts=2):
        """Host-memory runNlBayes.  Returns (basic, final, groups_ratio)."""
        noisy = np.ascontiguousarray(noisy, np.float32)
        F, Cc, H, W = noisy.shape
        b = np.zeros_like(noisy) if basic is None else np.ascontiguousarray(basic, np.float32).copy()
        fin = np.zeros_like(noisy) if p2.sizePatch else None
        gr = (C.c_float * 2)()
        ff = bf = None
        if fflow is not None:
            ff = np.ascontiguousarray(fflow, np.float32)
            bf = np.ascontiguousarray(bflow, np.float32)
        self._check(self._L.bm4dof_run(
            self.h, C.byref(p1), C.byref(p2), noisy.ctypes.data, W, H, F, Cc,
            ff.ctypes.data if ff is not None else None, bf.ctypes.data if bf is not None else None,
            b.ctypes.data, fin.ctypes.data if fin is not None else None, nparts, gr))
        return b, fin, (gr[0], gr[1])

    def run_dev(self, p1, p2, noisy_ptr, shape, fflow_ptr, bflow_ptr, basic_ptr, final_ptr,
                nparts=2):
        """Device-memory runNlBayes on raw device pointers (e.g. torch
        tensors' data_ptr()).  Returns groups_ratio."""
        F, Cc, H, W = shape
        gr = (C.c_float * 2)()
        self._check(self._L.bm4dof_run_dev(self.h, C.byref(p1), C.byref(p2), noisy_ptr, W, H, F,
                                           Cc, fflow_ptr, bflow_ptr, basic_ptr, final_ptr,
                                           nparts, gr))
        return gr[0], gr[1]

    def set_option(self, key, value):
        """bm4dof_ctx_set_option: select between kernels of equal output."""
        self._check(self._L.bm4dof_ctx_set_option(self.h, key.encode(), int(value)))

    def timings(self):
        ms = (C.c_float * 4096)()
        names = (C.c_char_p * 4096)()
        n = self._L.bm4dof_last_timings(self.h, ms, names, 4096)
        return [(names[i].decode(), ms[i]) for i in range(n)]

    def stats(self):
        """Statistics of the last run: dict with the ordered aggregation's
        pixel blocks / banded blocks / block shape and the groups and group
        rows of each step, and the executed TV-L1 iterations x pixels of the
        last flow computation."""
        v = (C.c_int64 * 10)()
        n = self._L.bm4dof_last_stats(self.h, v, 10)
        keys = ("agg_blocks", "agg_banded_blocks", "agg_bx", "agg_by", "agg_bt",
                "groups1", "rows1", "groups2", "rows2", "tvl1_pixel_iters")
        return {keys[i]: int(v[i]) for i in range(n)}

    def bm(self, p, image, fflow=None, bflow=None, nparts=2):
        """Neighbour lists at every reference position (bm4dof_bm)."""
        image = np.ascontiguousarray(image, np.float32)
        F, Cc, H, W = image.shape
        n = C.c_int64()
        self._check(self._L.bm4dof_bm_count(self.h, C.byref(p), W, H, F, Cc, nparts, C.byref(n)))
        G = n.value
        npm = max(1, p.nSimilarPatches)
        idx = np.zeros(G * npm, np.uint32)
        nsim = np.zeros(G, np.uint32)
        ref = np.zeros(G, np.uint32)
        tile = np.zeros(G, np.uint32)
        ff = bf = None
        if fflow is not None:
            ff = np.ascontiguousarray(fflow, np.float32)
            bf = np.ascontiguousarray(bflow, np.float32)
        self._check(self._L.bm4dof_bm(self.h, C.byref(p), image.ctypes.data, W, H, F, Cc,
                                      ff.ctypes.data if ff is not None else None,
                                      bf.ctypes.data if bf is not None else None, nparts,
                                      idx.ctypes.data, nsim.ctypes.data, ref.ctypes.data,
                                      tile.ctypes.data))
        return idx.reshape(G, npm), nsim, ref, tile

    # ------------------------------------------------------------- batches

    def batch(self, seqs, p1s, p2s, nparts=2, graph=True, autotune=False):
        """bm4dof_batch_create over Seq descriptors; returns a Batch."""
        n = len(seqs)
        arr = (Seq * n)(*seqs)
        a1 = (Params * n)(*p1s)
        a2 = (Params * n)(*p2s)
        h = C.c_void_p()
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_AUTOTUNE if autotune else 0)
        self._check(self._L.bm4dof_batch_create(self.h, arr, a1, a2, n, nparts, flags, C.byref(h)))
        return Batch(self, h)

    def autotune(self, p1, p2, noisy_ptr, shape, ff_ptr, bf_ptr, basic_ptr, final_ptr, nparts=2):
        F, Cc, H, W = shape
        rep = C.create_string_buffer(4096)
        self._check(self._L.bm4dof_autotune(self.h, C.byref(p1), C.byref(p2), noisy_ptr, W, H, F, ff_ptr,
                                            bf_ptr, basic_ptr, final_ptr, nparts, rep, 4096))
        return rep.value.decode()

    # ------------------------------------------------------------- learned pass

    def mlp_set(self, theta):
        theta = np.ascontiguousarray(theta, np.float32)
        assert theta.size == MLP_NPARAMS
        self._check(self._L.bm4dof_mlp_set(self.h, theta.ctypes.data))

    def mlp_get(self):
        theta = np.zeros(MLP_NPARAMS, np.float32)
        self._check(self._L.bm4dof_mlp_get(self.h, theta.ctypes.data))
        return theta

    def mlp_forward(self, x):
        """Gains of feature rows x (B, 66) on the matrix cores -> (B, 64)."""
        x = np.ascontiguousarray(x, np.float32)
        g = np.zeros((x.shape[0], 64), np.float32)
        self._check(self._L.bm4dof_mlp_forward(self.h, x.ctypes.data, x.shape[0], g.ctypes.data, 0))
        return g

    def mlp_grad(self, x, t):
        """(loss, gradient) of one batch by the training kernels."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_NPARAMS, np.float32)
        loss = C.c_float()
        self._check(self._L.bm4dof_mlp_grad(self.h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                            C.byref(loss), g.ctypes.data))
        return loss.value, g

    def mlp_train_dev(self, x_ptr, t_ptr, rows, batch, steps, lr, seed=1):
        """Adam steps on a device data set; returns the per-step losses."""
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(self._L.bm4dof_mlp_train(self.h, x_ptr, t_ptr, rows, batch, steps, lr, seed,
                                             losses.ctypes.data))
        return losses[:steps]

    def learned_pairs_dev(self, p1, noisy_ptr, basic_ptr, shape, ff_ptr, bf_ptr, nparts, stride, beta2,
                          x_ptr, t_ptr, flag_ptr, max_rows):
        F, Cc, H, W = shape
        rows = C.c_int64()
        self._check(self._L.bm4dof_learned_pairs(self.h, C.byref(p1), noisy_ptr, basic_ptr, W, H, F,
                                                 ff_ptr, bf_ptr, nparts, stride, beta2, x_ptr, t_ptr,
                                                 flag_ptr, max_rows, C.byref(rows)))
        return rows.value

    def run_learned(self, p1, p2, noisy, fflow=None, bflow=None, nparts=2, gvar_thr=-1.0):
        """Learned single pass on host arrays -> (output, flagged fraction)."""
        noisy = np.ascontiguousarray(noisy, np.float32)
        F, Cc, H, W = noisy.shape
        out = np.zeros_like(noisy)
        ff = bf = None
        if fflow is not None:
            ff = np.ascontiguousarray(fflow, np.float32)
            bf = np.ascontiguousarray(bflow, np.float32)
        frac = C.c_float()
        self._check(self._L.bm4dof_run_learned(
            self.h, C.byref(p1), C.byref(p2) if p2 is not None else None, noisy.ctypes.data, W, H, F, Cc,
            ff.ctypes.data if ff is not None else None, bf.ctypes.data if bf is not None else None,
            out.ctypes.data, nparts, gvar_thr, C.byref(frac)))
        return out, frac.value

    def run_learned_dev(self, p1, p2, noisy_ptr, shape, ff_ptr, bf_ptr, out_ptr, nparts=2, gvar_thr=-1.0):
        F, Cc, H, W = shape
        frac = C.c_float()
        self._check(self._L.bm4dof_run_learned_dev(
            self.h, C.byref(p1), C.byref(p2) if p2 is not None else None, noisy_ptr, W, H, F, Cc, ff_ptr,
            bf_ptr, out_ptr, nparts, gvar_thr, C.byref(frac)))
        return frac.value

    # ------------------------------------------------------------- TV-L1

    def tvl1(self, I0, I1, tau=0.25, lam=0.15, theta=0.3, nscales=5, fscale=0, zfactor=0.5,
             warps=10, eps=0.01):
        """Flow of one frame pair on host arrays -> (u, v, iters[nscales, warps])."""
        I0 = np.ascontiguousarray(I0, np.float32)
        I1 = np.ascontiguousarray(I1, np.float32)
        H, W = I0.shape
        u = np.zeros((H, W), np.float32)
        v = np.zeros((H, W), np.float32)
        it = np.zeros(nscales * warps, np.int32)
        prm = TVL1Params(tau, lam, theta, nscales, fscale, zfactor, warps, eps, 0)
        self._check(self._L.bm4dof_tvl1(self.h, C.byref(prm), I0.ctypes.data, I1.ctypes.data,
                                        u.ctypes.data, v.ctypes.data, W, H, 0, it.ctypes.data))
        return u, v, it.reshape(nscales, warps)

    def tvl1_sequence(self, frames, tau=0.25, lam=0.15, theta=0.3, nscales=5, fscale=0,
                      zfactor=0.5, warps=10, eps=0.01):
        """Forward / backward flow of a (F, H, W) gray sequence ->
        (fflow, bflow), each (F, 2, H, W)."""
        frames = np.ascontiguousarray(frames, np.float32)
        F, H, W = frames.shape
        ff = np.zeros((F, 2, H, W), np.float32)
        bf = np.zeros((F, 2, H, W), np.float32)
        prm = TVL1Params(tau, lam, theta, nscales, fscale, zfactor, warps, eps, 0)
        self._check(self._L.bm4dof_tvl1_sequence(self.h, C.byref(prm), frames.ctypes.data, W, H,
                                                 F, ff.ctypes.data, bf.ctypes.data, 0))
        return ff, bf

    def tvl1_sequence_dev(self, frames_ptr, W, H, F, fflow_ptr, bflow_ptr, tau=0.25, lam=0.15,
                          theta=0.3, nscales=5, fscale=0, zfactor=0.5, warps=10, eps=0.01):
        """bm4dof_tvl1_sequence on device pointers (frames: F*H*W floats;
        flows: F*2*H*W floats each)."""
        prm = TVL1Params(tau, lam, theta, nscales, fscale, zfactor, warps, eps, 0)
        self._check(self._L.bm4dof_tvl1_sequence(self.h, C.byref(prm), frames_ptr, W, H, F,
                                                 fflow_ptr, bflow_ptr, 1))


class Batch:
    """A created batch (bm4dof_batch); run() replays it."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def run(self):
        self.ctx._check(self.ctx._L.bm4dof_batch_run(self.h))
        return float(self.ctx._L.bm4dof_batch_last_ms(self.h))

    def report(self):
        return self.ctx._L.bm4dof_batch_report(self.h).decode()

    def stats(self):
        """How the last run ran: path (0 enqueued, 1 graph replayed, 2 graph
        captured again), reruns (sequences run again for the mask scan),
        captures (so far)."""
        v = (C.c_int64 * 3)()
        n = self.ctx._L.bm4dof_batch_stats(self.h, v, 3)
        return dict(zip(("path", "reruns", "captures"), (int(v[i]) for i in range(n))))

    def close(self):
        if self.h:
            self.ctx._L.bm4dof_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- image files

def io_lib():
    """Load libbm4dof_io.so (host code: PNG, TIFF and .flo files)."""
    global _IO
    if _IO is not None:
        return _IO
    if not os.path.exists(IO_PATH):
        raise ImportError(f"{IO_PATH} is missing: build it with `make -C {_HERE}`")
    L = C.CDLL(IO_PATH)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int)
    L.bm4dof_io_read.argtypes = [C.c_char_p, C.POINTER(vp), ip, ip, ip]
    L.bm4dof_io_read_gray.argtypes = [C.c_char_p, C.POINTER(vp), ip, ip]
    L.bm4dof_io_write.argtypes = [C.c_char_p, vp, C.c_int, C.c_int, C.c_int]
    L.bm4dof_io_free.argtypes = [vp]
    L.bm4dof_io_free.restype = None
    L.bm4dof_io_error.restype = C.c_char_p
    _IO = L
    return L


def _take(L, ptr, n):
    out = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(n,)).copy()
    L.bm4dof_io_free(ptr)
    return out


def read_image(path):
    """A PNG, TIFF or .flo file as a (C, H, W) float32 array."""
    L = io_lib()
    ptr, w, h, c = C.c_void_p(), C.c_int(), C.c_int(), C.c_int()
    if L.bm4dof_io_read(os.fsencode(path), C.byref(ptr), C.byref(w), C.byref(h), C.byref(c)):
        raise OSError(L.bm4dof_io_error().decode())
    return _take(L, ptr, c.value * h.value * w.value).reshape(c.value, h.value, w.value)


def read_gray(path):
    """One gray (H, W) channel, read the way tvl1flow reads its frames."""
    L = io_lib()
    ptr, w, h = C.c_void_p(), C.c_int(), C.c_int()
    if L.bm4dof_io_read_gray(os.fsencode(path), C.byref(ptr), C.byref(w), C.byref(h)):
        raise OSError(L.bm4dof_io_error().decode())
    return _take(L, ptr, h.value * w.value).reshape(h.value, w.value)


def write_image(path, img):
    """Write a (C, H, W) or (H, W) array; the format follows the extension."""
    a = np.ascontiguousarray(img, np.float32)
    if a.ndim == 2:
        a = a[None]
    L = io_lib()
    if L.bm4dof_io_write(os.fsencode(path), a.ctypes.data, a.shape[2], a.shape[1], a.shape[0]):
        raise OSError(L.bm4dof_io_error().decode())
