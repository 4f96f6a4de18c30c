"""wsbb_amd: MI355X-native flowshop branch-and-bound bounding engine.

Python face of the C ABI in include/wsbb_amd.h (ctypes, numpy host buffers).
The compute path is the HIP library wsbb_amd/libwsbb_amd.so; importing this
module fails loudly if it has not been built (python wsbb_amd/build.py).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# WSBB_AMD_LIB names another build of the library (side-by-side timing of
# two builds in one process tree); unset, the in-tree build is loaded
LIB_PATH = os.environ.get("WSBB_AMD_LIB") or os.path.join(_HERE, "libwsbb_amd.so")

LB2, LB1, LEARNED = 0, 1, 2
NSTATS = 12
STAT_NAMES = ("decomposed", "bounded", "leaves", "pruned_pop", "pruned_insert",
              "improvements", "max_pool", "rounds", "subset_bounded", "fallbacks",
              "pairs_evaluated", "work_ops")
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EOVERFLOW", -5: "EBUDGET",
          -6: "ENODEV"}

_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


# wsbb_trace_fn: (user, round, np, permT, l1, l2, cost, nc, desc, bound, inc)
_TRACE_FN = ctypes.CFUNCTYPE(None, _vp, ctypes.c_int64, ctypes.c_int, _u8p, _i32p, _i32p, _i32p,
                             ctypes.c_int, _i32p, _i32p, ctypes.c_int32)


class WsbbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python wsbb_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "wsbb_version": ([], ctypes.c_int),
        "wsbb_last_error": ([], ctypes.c_char_p),
        "wsbb_device_count": ([_i32p], ctypes.c_int),
        "wsbb_taillard": ([ctypes.c_int, _i32p, _i32p, _i32p], ctypes.c_int),
        "wsbb_generate": ([ctypes.c_int, ctypes.c_int, ctypes.c_long, _i32p], ctypes.c_int),
        "wsbb_generate_correlated": ([ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _i32p], ctypes.c_int),
        "wsbb_generate_machine_correlated": ([ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _i32p],
                                             ctypes.c_int),
        "wsbb_random_nodes": ([ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _u8p, _i32p, _i32p], ctypes.c_int),
        "wsbb_precompute": ([_i32p, ctypes.c_int, ctypes.c_int] + [_i32p] * 6, ctypes.c_int),
        "wsbb_create": ([_i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
        "wsbb_destroy": ([_vp], ctypes.c_int),
        "wsbb_info": ([_vp, _i32p, _i32p, _i32p, _i32p], ctypes.c_int),
        "wsbb_bound_batch": ([_vp, _u8p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p], ctypes.c_int),
        "wsbb_bound_batch_ex": ([_vp, _u8p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 _i32p, _i32p, _i32p, _u32p, _f32p], ctypes.c_int),
        "wsbb_bound_device": ([_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
        "wsbb_bound_batch_noalloc": ([_vp, _u8p, _i32p, _i32p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _i32p], ctypes.c_int),
        "wsbb_bound_tiles": ([_vp, _u8p, _i32p, _i32p, ctypes.c_int, _u32p, ctypes.c_int,
                              ctypes.c_int32, _i32p, _i64p], ctypes.c_int),
        "wsbb_malloc": ([_vp, ctypes.c_size_t, ctypes.POINTER(_vp)], ctypes.c_int),
        "wsbb_free": ([_vp, _vp], ctypes.c_int),
        "wsbb_h2d": ([_vp, _vp, _vp, ctypes.c_size_t], ctypes.c_int),
        "wsbb_d2h": ([_vp, _vp, _vp, ctypes.c_size_t], ctypes.c_int),
        "wsbb_synchronize": ([_vp], ctypes.c_int),
        "wsbb_timer_start": ([_vp], ctypes.c_int),
        "wsbb_timer_stop": ([_vp, _f32p], ctypes.c_int),
        "wsbb_set_launch": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "wsbb_autotune": ([_vp, ctypes.c_int, ctypes.c_int, _i32p, _f32p, _i32p], ctypes.c_int),
        "wsbb_solve": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int32, ctypes.c_int64,
                        _i32p, _i32p, _i64p], ctypes.c_int),
        "wsbb_search_profile": ([_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _i64p],
                                ctypes.c_int),
        "wsbb_set_pool_capacity": ([_vp, ctypes.c_int], ctypes.c_int),
        "wsbb_search_options": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "wsbb_search_trace": ([_vp, _TRACE_FN, _vp], ctypes.c_int),
        "wsbb_pool_open": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "wsbb_pool_seed_root": ([_vp], ctypes.c_int),
        "wsbb_pool_push": ([_vp, _u8p, _i32p, _i32p, _i32p, ctypes.c_int], ctypes.c_int),
        "wsbb_pool_pop": ([_vp, ctypes.c_int, ctypes.c_int, _u8p, _i32p, _i32p, _i32p, _i32p],
                          ctypes.c_int),
        "wsbb_pool_size": ([_vp, _i32p], ctypes.c_int),
        "wsbb_incumbent_get": ([_vp, _i32p, _i32p], ctypes.c_int),
        "wsbb_incumbent_set": ([_vp, ctypes.c_int32, _i32p], ctypes.c_int),
        "wsbb_pool_run": ([_vp, ctypes.c_int, ctypes.c_int64, _i32p, _i64p], ctypes.c_int),
        "wsbb_set_int16": ([_vp, ctypes.c_int], ctypes.c_int),
        "wsbb_multi_create": ([ctypes.c_int, ctypes.POINTER(_vp), _i32p, _i32p, ctypes.c_int,
                               ctypes.POINTER(_vp)], ctypes.c_int),
        "wsbb_multi_destroy": ([_vp], ctypes.c_int),
        "wsbb_multi_context": ([_vp, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
        "wsbb_multi_load": ([_vp, _u8p, _i32p, _i32p, _i32p, ctypes.c_int], ctypes.c_int),
        "wsbb_multi_run": ([_vp, ctypes.c_int, _i32p, ctypes.c_int, _u32p, ctypes.c_int, _i32p,
                            _i64p, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "wsbb_search_state": ([_vp, _i32p, _i64p, _i32p, _i32p, _u8p, _i32p, _i32p, _i32p,
                               ctypes.c_int], ctypes.c_int),
        "wsbb_solve_resume": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                               _u8p, _i32p, _i32p, _i32p, ctypes.c_int, ctypes.c_int32, _i64p,
                               _i32p, _i32p, _i32p, _i64p], ctypes.c_int),
        "wsbb_copy_device": ([_vp, _vp, _vp, ctypes.c_size_t], ctypes.c_int),
        "wsbb_valu_probe": ([_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_size_t,
                             ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "wsbb_selector_init": ([_vp, ctypes.c_uint64], ctypes.c_int),
        "wsbb_selector_params": ([_vp, ctypes.c_int, _f32p, _f32p, _f32p, _f32p], ctypes.c_int),
        "wsbb_selector_config": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "wsbb_selector_log": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "wsbb_selector_records": ([_vp, _f32p, _i32p, _i32p], ctypes.c_int),
        "wsbb_selector_set_records": ([_vp, _f32p, _i32p, ctypes.c_int], ctypes.c_int),
        "wsbb_selector_grad": ([_vp, _i32p, ctypes.c_int, _f32p, _f32p, _f32p, _f32p, _f32p],
                               ctypes.c_int),
        "wsbb_selector_train": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                 _f32p], ctypes.c_int),
        "wsbb_selector_infer": ([_vp, _f32p, ctypes.c_int, _f32p, _u32p], ctypes.c_int),
        "wsbb_selector_parents": ([_vp, ctypes.c_int], ctypes.c_int),
        "wsbb_selector_infer_parents": ([_vp, _f32p, ctypes.c_int, _i32p, _f32p, _u32p],
                                        ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


lib = _load()


def _p(a, t=_i32p):
    return None if a is None else a.ctypes.data_as(t)


def _check(rc, what=""):
    if rc != 0:
        raise WsbbError(rc, (what + ": " if what else "") + lib.wsbb_last_error().decode())
    return rc


def device_count():
    c = ctypes.c_int32()
    lib.wsbb_device_count(ctypes.byref(c))
    return c.value


def taillard(tid):
    """Taillard instance `tid`: returns p[m, n] int32 (machine-major)."""
    n, m = ctypes.c_int32(), ctypes.c_int32()
    _check(lib.wsbb_taillard(tid, ctypes.byref(n), ctypes.byref(m), None), "taillard")
    p = np.zeros((m.value, n.value), np.int32)
    _check(lib.wsbb_taillard(tid, None, None, _p(p)), "taillard")
    return p


def generate(n, m, seed):
    p = np.zeros((m, n), np.int32)
    _check(lib.wsbb_generate(n, m, seed, _p(p)))
    return p


def generate_correlated(n, m, seed):
    p = np.zeros((m, n), np.int32)
    _check(lib.wsbb_generate_correlated(n, m, seed, _p(p)))
    return p


def generate_machine_correlated(n, m, seed):
    p = np.zeros((m, n), np.int32)
    _check(lib.wsbb_generate_machine_correlated(n, m, seed, _p(p)))
    return p


def random_nodes(n, N, seed0=0x5EED):
    permT = np.zeros((n, N), np.uint8)
    l1 = np.zeros(N, np.int32)
    l2 = np.zeros(N, np.int32)
    _check(lib.wsbb_random_nodes(n, N, seed0, _p(permT, _u8p), _p(l1), _p(l2)))
    return permT, l1, l2


def precompute(p):
    p = np.ascontiguousarray(p, np.int32)
    m, n = p.shape
    P = m * (m - 1) // 2
    out = dict(ma=np.zeros(P, np.int32), mb=np.zeros(P, np.int32),
               lag=np.zeros((P, n), np.int32), johnson=np.zeros((P, n), np.int32),
               R0=np.zeros(m, np.int32), Q0=np.zeros(m, np.int32))
    _check(lib.wsbb_precompute(_p(p), n, m, *[_p(out[k]) for k in
                                              ("ma", "mb", "lag", "johnson", "R0", "Q0")]))
    return out


class Engine:
    """One device context: an instance's tables in HBM plus a HIP stream."""

    def __init__(self, p, device=0, _handle=None):
        self.p = np.ascontiguousarray(p, np.int32)
        self.m, self.n = self.p.shape
        self._owned = _handle is None
        if _handle is None:
            h = _vp()
            _check(lib.wsbb_create(_p(self.p), self.n, self.m, device, ctypes.byref(h)), "create")
        else:
            h = _handle
        self.h = h
        P, cus = ctypes.c_int32(), ctypes.c_int32()
        lib.wsbb_info(self.h, None, None, ctypes.byref(P), ctypes.byref(cus))
        self.P, self.num_cus = P.value, cus.value

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                lib.wsbb_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- batch bounding (host buffers) ----
    def bound(self, permT, l1, l2, kind=LB2, best=-1, argmax=False, pairv=False, sel=None,
              feat=False):
        permT = np.ascontiguousarray(permT, np.uint8)
        l1 = np.ascontiguousarray(l1, np.int32)
        l2 = np.ascontiguousarray(l2, np.int32)
        N = l1.shape[0]
        assert permT.shape == (self.n, N) and l2.shape == (N,)
        out = np.zeros(N, np.int32)
        am = np.zeros(N, np.int32) if argmax else None
        pv = np.zeros((self.P, N), np.int32) if pairv else None
        ft = np.zeros((5 * self.m, N), np.float32) if feat else None
        if sel is not None:
            sel = np.ascontiguousarray(sel, np.uint32)
            assert sel.shape == ((self.P + 31) // 32, N)
        _check(lib.wsbb_bound_batch_ex(self.h, _p(permT, _u8p), _p(l1), _p(l2), N, kind, best,
                                       _p(out), _p(am), _p(pv), _p(sel, _u32p), _p(ft, _f32p)),
               "bound")
        if not (argmax or pairv or feat):
            return out
        res = {"bound": out}
        if argmax:
            res["argmax"] = am
        if pairv:
            res["pairv"] = pv
        if feat:
            res["feat"] = ft
        return res

    def bound_noalloc(self, permT, l1, l2, kind=LB2, best=-1):
        """wsbb_bound_batch_noalloc: the batch bound staged through the
        library's fixed per-GPU buffer (no device allocation)."""
        permT = np.ascontiguousarray(permT, np.uint8)
        l1 = np.ascontiguousarray(l1, np.int32)
        l2 = np.ascontiguousarray(l2, np.int32)
        N = l1.shape[0]
        assert permT.shape == (self.n, N) and l2.shape == (N,)
        out = np.zeros(N, np.int32)
        _check(lib.wsbb_bound_batch_noalloc(self.h, _p(permT, _u8p), _p(l1), _p(l2), N, kind,
                                            best, _p(out)), "bound_noalloc")
        return out

    def bound_tiles(self, permT, l1, l2, tile_sel, margin=-1, inc=0):
        """LB2 over a pair subset per 64-node tile (+ full-LB2 fallback)."""
        permT = np.ascontiguousarray(permT, np.uint8)
        l1 = np.ascontiguousarray(l1, np.int32)
        l2 = np.ascontiguousarray(l2, np.int32)
        tile_sel = np.ascontiguousarray(tile_sel, np.uint32)
        N = l1.shape[0]
        assert tile_sel.shape == ((N + 63) // 64, (self.P + 31) // 32)
        out = np.zeros(N, np.int32)
        st = np.zeros(3, np.int64)
        _check(lib.wsbb_bound_tiles(self.h, _p(permT, _u8p), _p(l1), _p(l2), N,
                                    _p(tile_sel, _u32p), margin, inc, _p(out), _p(st, _i64p)),
               "bound_tiles")
        return out, dict(nodes=int(st[0]), fallbacks=int(st[1]), pair_evals=int(st[2]))

    # ---- device memory ----
    def malloc(self, nbytes):
        d = _vp()
        _check(lib.wsbb_malloc(self.h, nbytes, ctypes.byref(d)), "malloc")
        return d.value

    def free(self, d):
        _check(lib.wsbb_free(self.h, d), "free")

    def h2d(self, d, arr):
        arr = np.ascontiguousarray(arr)
        _check(lib.wsbb_h2d(self.h, d, arr.ctypes.data, arr.nbytes), "h2d")

    def d2h(self, arr, d):
        _check(lib.wsbb_d2h(self.h, arr.ctypes.data, d, arr.nbytes), "d2h")
        return arr

    def bound_device(self, d_perm, d_l1, d_l2, N, d_out, kind=LB2):
        _check(lib.wsbb_bound_device(self.h, d_perm, d_l1, d_l2, N, kind, d_out), "bound_device")

    def copy_device(self, d_dst, d_src, nbytes):
        _check(lib.wsbb_copy_device(self.h, d_dst, d_src, nbytes), "copy_device")

    def valu_probe(self, kind, iters, d_out, nbytes):
        """Launches the integer VALU probe; returns the lane-ops it executes."""
        ops = ctypes.c_double()
        _check(lib.wsbb_valu_probe(self.h, kind, iters, d_out, nbytes, ctypes.byref(ops)),
               "valu_probe")
        return ops.value

    def synchronize(self):
        _check(lib.wsbb_synchronize(self.h), "synchronize")

    def timer_start(self):
        _check(lib.wsbb_timer_start(self.h), "timer_start")

    def timer_stop(self):
        ms = ctypes.c_float()
        _check(lib.wsbb_timer_stop(self.h, ctypes.byref(ms)), "timer_stop")
        return ms.value

    def set_launch(self, groups=0, grid_blocks=0):
        _check(lib.wsbb_set_launch(self.h, groups, grid_blocks), "set_launch")

    def autotune(self, N, reps=3):
        shapes = np.zeros((16, 2), np.int32)
        ms = np.zeros(16, np.float32)
        cnt = ctypes.c_int32()
        _check(lib.wsbb_autotune(self.h, N, reps, _p(shapes), _p(ms, _f32p), ctypes.byref(cnt)),
               "autotune")
        k = cnt.value
        return [(int(g), int(b), float(t)) for (g, b), t in zip(shapes[:k], ms[:k])]

    # ---- search ----
    def search_profile(self, enable=-1):
        """enable = k > 0: time the bound launch of every k-th round of later
        solves (0 off, -1 keep); returns (summed ms, timed launches) of the
        last solve."""
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        _check(lib.wsbb_search_profile(self.h, enable, ctypes.byref(ms), ctypes.byref(cnt)),
               "search_profile")
        return ms.value, cnt.value

    def set_pool_capacity(self, nodes):
        _check(lib.wsbb_set_pool_capacity(self.h, nodes), "set_pool_capacity")

    def set_int16(self, mode):
        """Two nodes per lane in int16 halves for plain LB2: 0 never, 1 for
        large batches (default), 2 whenever the instance fits."""
        _check(lib.wsbb_set_int16(self.h, int(mode)), "set_int16")

    def search_options(self, use_graph=-1, rounds_per_sync=-1):
        """hipGraph replay of the search rounds (1/0) and rounds per host
        check; -1 keeps a setting."""
        _check(lib.wsbb_search_options(self.h, use_graph, rounds_per_sync), "search_options")

    # ---- learned pair selector ----
    HIDDEN = 128

    def selector_init(self, seed=1):
        _check(lib.wsbb_selector_init(self.h, seed), "selector_init")

    def selector_params(self, params=None):
        """Download the FP32 masters, or upload `params` (dict W1, b1, W2, b2)."""
        F, H, P = 5 * self.m, self.HIDDEN, self.P
        if params is None:
            out = dict(W1=np.zeros((F, H), np.float32), b1=np.zeros(H, np.float32),
                       W2=np.zeros((H, P), np.float32), b2=np.zeros(P, np.float32))
            _check(lib.wsbb_selector_params(self.h, 0, *[_p(out[k], _f32p) for k in
                                                         ("W1", "b1", "W2", "b2")]), "params")
            return out
        arrs = [np.ascontiguousarray(params[k], np.float32) for k in ("W1", "b1", "W2", "b2")]
        _check(lib.wsbb_selector_params(self.h, 1, *[_p(a, _f32p) for a in arrs]), "params")

    def selector_config(self, topk=16, margin=0, parents=None):
        """parents: True = the network scores parents (records hold the
        parent's features with each child's label), False = children; None
        keeps the current setting."""
        _check(lib.wsbb_selector_config(self.h, topk, margin), "selector_config")
        if parents is not None:
            _check(lib.wsbb_selector_parents(self.h, int(bool(parents))), "selector_parents")

    def selector_log(self, stride, capacity):
        _check(lib.wsbb_selector_log(self.h, stride, capacity), "selector_log")

    def selector_records(self):
        cnt = ctypes.c_int32()
        _check(lib.wsbb_selector_records(self.h, None, None, ctypes.byref(cnt)), "records")
        c = cnt.value
        feat = np.zeros((5 * self.m, c), np.float32)
        lab = np.zeros(c, np.int32)
        if c:
            _check(lib.wsbb_selector_records(self.h, _p(feat, _f32p), _p(lab), ctypes.byref(cnt)),
                   "records")
        return feat, lab

    def selector_set_records(self, feat, label):
        feat = np.ascontiguousarray(feat, np.float32)
        label = np.ascontiguousarray(label, np.int32)
        _check(lib.wsbb_selector_set_records(self.h, _p(feat, _f32p), _p(label), label.shape[0]),
               "set_records")

    def selector_grad(self, idx):
        idx = np.ascontiguousarray(idx, np.int32)
        F, H, P = 5 * self.m, self.HIDDEN, self.P
        g = dict(W1=np.zeros((F, H), np.float32), b1=np.zeros(H, np.float32),
                 W2=np.zeros((H, P), np.float32), b2=np.zeros(P, np.float32))
        loss = ctypes.c_float()
        _check(lib.wsbb_selector_grad(self.h, _p(idx), idx.shape[0], ctypes.byref(loss),
                                      *[_p(g[k], _f32p) for k in ("W1", "b1", "W2", "b2")]),
               "grad")
        return loss.value, g

    def selector_train(self, epochs=5, batch=512, lr=1e-3, seed=1):
        losses = np.zeros(epochs, np.float32)
        _check(lib.wsbb_selector_train(self.h, epochs, batch, lr, seed, _p(losses, _f32p)),
               "train")
        return losses

    def selector_infer(self, feat):
        feat = np.ascontiguousarray(feat, np.float32)
        N = feat.shape[1]
        scores = np.zeros((self.P, N), np.float32)
        tsel = np.zeros(((N + 63) // 64, (self.P + 31) // 32), np.uint32)
        _check(lib.wsbb_selector_infer(self.h, _p(feat, _f32p), N, _p(scores, _f32p),
                                       _p(tsel, _u32p)), "infer")
        return scores, tsel

    def selector_infer_parents(self, feat, counts):
        """features [5m][Np] of parents with counts[p] consecutive children:
        (probs [Np][P], tile masks [ceil(sum counts / 64)][ceil(P/32)])"""
        feat = np.ascontiguousarray(feat, np.float32)
        counts = np.ascontiguousarray(counts, np.int32)
        Np = feat.shape[1]
        assert counts.shape == (Np,)
        probs = np.zeros((Np, self.P), np.float32)
        tsel = np.zeros(((int(counts.sum()) + 63) // 64, (self.P + 31) // 32), np.uint32)
        _check(lib.wsbb_selector_infer_parents(self.h, _p(feat, _f32p), Np, _p(counts),
                                               _p(probs, _f32p), _p(tsel, _u32p)),
               "infer_parents")
        return probs, tsel

    def search_trace(self, on=True):
        """Record every round of later solves (wsbb_search_trace): returns
        the list the rounds are appended to, one dict per round with the
        parents in pop order (permT [n, np], l1, l2, cost), the children in
        generation order (desc, bound) and the incumbent after the round.
        on=False turns it off."""
        if not on:
            _check(lib.wsbb_search_trace(self.h, _TRACE_FN(), None), "search_trace")
            self._trace_cb = None
            return None
        rounds = []
        n = self.n

        def cb(_user, rnd, np_, permT, l1, l2, cost, nc, desc, bound, inc):
            def arr(ptr, cnt, dt):
                return np.ctypeslib.as_array(ptr, shape=(cnt,)).astype(dt, copy=True) \
                    if cnt else np.zeros(0, dt)
            rounds.append(dict(
                round=int(rnd),
                permT=arr(permT, n * np_, np.uint8).reshape(n, np_),
                l1=arr(l1, np_, np.int32), l2=arr(l2, np_, np.int32),
                cost=arr(cost, np_, np.int32), desc=arr(desc, nc, np.int32),
                bound=arr(bound, nc, np.int32), inc=int(inc)))

        self._trace_cb = _TRACE_FN(cb)
        _check(lib.wsbb_search_trace(self.h, self._trace_cb, None), "search_trace")
        return rounds

    # ---- live node pool (wsbb_pool_*) ----
    def pool_open(self, kind=LB2, pool_size=4096, max_batches=1):
        _check(lib.wsbb_pool_open(self.h, kind, pool_size, max_batches), "pool_open")

    def pool_seed_root(self):
        _check(lib.wsbb_pool_seed_root(self.h), "pool_seed_root")

    def pool_push(self, nodes):
        """nodes: dict permT [n, count], l1, l2, cost (bottom-first)."""
        permT = np.ascontiguousarray(nodes["permT"], np.uint8)
        l1, l2, cost = (np.ascontiguousarray(nodes[k], np.int32) for k in ("l1", "l2", "cost"))
        c = l1.shape[0]
        assert permT.shape == (self.n, c) and l2.shape == cost.shape == (c,)
        _check(lib.wsbb_pool_push(self.h, _p(permT, _u8p), _p(l1), _p(l2), _p(cost), c),
               "pool_push")

    def pool_pop(self, max_nodes, from_bottom=False):
        permT = np.zeros((self.n, max(max_nodes, 1)), np.uint8)
        l1, l2, cost = (np.zeros(max(max_nodes, 1), np.int32) for _ in range(3))
        cnt = ctypes.c_int32()
        _check(lib.wsbb_pool_pop(self.h, max_nodes, int(from_bottom), _p(permT, _u8p), _p(l1),
                                 _p(l2), _p(cost), ctypes.byref(cnt)), "pool_pop")
        c = cnt.value
        # wsbb_pool_pop writes position-major rows of `c` entries
        flat = permT.reshape(-1)[: self.n * c].reshape(self.n, c).copy()
        return dict(permT=flat, l1=l1[:c].copy(), l2=l2[:c].copy(), cost=cost[:c].copy())

    def pool_size(self):
        sz = ctypes.c_int32()
        _check(lib.wsbb_pool_size(self.h, ctypes.byref(sz)), "pool_size")
        return sz.value

    def incumbent(self):
        inc = ctypes.c_int32()
        perm = np.zeros(self.n, np.int32)
        _check(lib.wsbb_incumbent_get(self.h, ctypes.byref(inc), _p(perm)), "incumbent_get")
        return inc.value, perm

    def set_incumbent(self, inc, perm=None):
        pp = None if perm is None else np.ascontiguousarray(perm, np.int32)
        _check(lib.wsbb_incumbent_set(self.h, int(inc), _p(pp)), "incumbent_set")

    def pool_run(self, rounds=0, budget=0):
        """Runs rounds of the open session; returns (code, pool size,
        counters) where code is 0, or -5 / -4 for a budget / overflow stop."""
        sz = ctypes.c_int32()
        st = np.zeros(NSTATS, np.int64)
        rc = lib.wsbb_pool_run(self.h, rounds, budget, ctypes.byref(sz), _p(st, _i64p))
        if rc not in (0, -4, -5):
            _check(rc, "pool_run")
        return rc, sz.value, dict(zip(STAT_NAMES, st.tolist()))

    def solve(self, kind=LB2, pool_size=4096, max_batches=1, ub0=2**31 - 1, budget=0):
        best = ctypes.c_int32()
        bp = np.zeros(self.n, np.int32)
        st = np.zeros(NSTATS, np.int64)
        rc = lib.wsbb_solve(self.h, kind, pool_size, max_batches, ub0, budget,
                            ctypes.byref(best), _p(bp), _p(st, _i64p))
        if rc not in (0, -5):
            _check(rc, "solve")
        return dict(best=best.value, perm=bp, stats=dict(zip(STAT_NAMES, st.tolist())),
                    proven=(rc == 0))

    def checkpoint(self):
        """Device state of the last solve: incumbent, counters, incumbent
        permutation and the node pool it stopped with."""
        inc, cnt = ctypes.c_int32(), ctypes.c_int32()
        st = np.zeros(NSTATS, np.int64)
        bp = np.zeros(self.n, np.int32)
        _check(lib.wsbb_search_state(self.h, ctypes.byref(inc), _p(st, _i64p), _p(bp),
                                     ctypes.byref(cnt), None, None, None, None, 0), "search_state")
        c = cnt.value
        permT = np.zeros((self.n, c), np.uint8)
        l1, l2, cost = (np.zeros(c, np.int32) for _ in range(3))
        _check(lib.wsbb_search_state(self.h, ctypes.byref(inc), _p(st, _i64p), _p(bp),
                                     ctypes.byref(cnt), _p(permT, _u8p), _p(l1), _p(l2), _p(cost),
                                     c), "search_state")
        return dict(inc=inc.value, stats=st, best_perm=bp, permT=permT, l1=l1, l2=l2, cost=cost)

    def resume(self, ckpt, kind=LB2, pool_size=4096, max_batches=1, budget=0):
        """Continue a search from a checkpoint() dict."""
        permT = np.ascontiguousarray(ckpt["permT"], np.uint8)
        l1, l2, cost = (np.ascontiguousarray(ckpt[k], np.int32) for k in ("l1", "l2", "cost"))
        st_in = np.ascontiguousarray(ckpt["stats"], np.int64)
        bp_in = np.ascontiguousarray(ckpt["best_perm"], np.int32)
        best = ctypes.c_int32()
        bp = np.zeros(self.n, np.int32)
        st = np.zeros(NSTATS, np.int64)
        rc = lib.wsbb_solve_resume(self.h, kind, pool_size, max_batches, budget, _p(permT, _u8p),
                                   _p(l1), _p(l2), _p(cost), l1.shape[0], int(ckpt["inc"]),
                                   _p(st_in, _i64p), _p(bp_in), ctypes.byref(best), _p(bp),
                                   _p(st, _i64p))
        if rc not in (0, -5):
            _check(rc, "resume")
        return dict(best=best.value, perm=bp, stats=dict(zip(STAT_NAMES, st.tolist())),
                    proven=(rc == 0))


class Multi:
    """Several instances whose nodes are bounded together, one launch per
    batch (wsbb_multi_*).  `instances` is a list of p[m, n] arrays."""

    def __init__(self, instances, device=0):
        self.ps = [np.ascontiguousarray(p, np.int32) for p in instances]
        k = len(self.ps)
        ptrs = (_vp * k)(*[ctypes.cast(p.ctypes.data, _vp) for p in self.ps])
        ns = np.array([p.shape[1] for p in self.ps], np.int32)
        ms = np.array([p.shape[0] for p in self.ps], np.int32)
        h = _vp()
        _check(lib.wsbb_multi_create(k, ptrs, _p(ns), _p(ms), device, ctypes.byref(h)),
               "multi_create")
        self.h = h
        self.maxn = int(ns.max())
        self.engines = []
        for i in range(k):
            c = _vp()
            _check(lib.wsbb_multi_context(self.h, i, ctypes.byref(c)), "multi_context")
            self.engines.append(Engine(self.ps[i], _handle=c))
        self.N = 0

    def close(self):
        if getattr(self, "h", None):
            for e in self.engines:
                e.close()
            lib.wsbb_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load(self, permT, l1, l2, offsets):
        """permT [max n][N] (rows past an instance's n unused), limits [N],
        offsets [count + 1]: instance i owns nodes offsets[i]..offsets[i+1]-1."""
        permT = np.ascontiguousarray(permT, np.uint8)
        l1 = np.ascontiguousarray(l1, np.int32)
        l2 = np.ascontiguousarray(l2, np.int32)
        offsets = np.ascontiguousarray(offsets, np.int32)
        self.N = l1.shape[0]
        assert permT.shape == (self.maxn, self.N)
        _check(lib.wsbb_multi_load(self.h, _p(permT, _u8p), _p(l1), _p(l2), _p(offsets), self.N),
               "multi_load")

    def run(self, kind=LB2, inc=None, margin=-1, tile_sel=None, reps=1):
        out = np.zeros(self.N, np.int32)
        st = np.zeros(3, np.int64)
        ms = ctypes.c_float()
        incp = None
        if inc is not None:
            inc = np.ascontiguousarray(inc, np.int32)
            incp = _p(inc)
        tsp = None
        if tile_sel is not None:
            tile_sel = np.ascontiguousarray(tile_sel, np.uint32)
            tsp = _p(tile_sel, _u32p)
        _check(lib.wsbb_multi_run(self.h, kind, incp, margin, tsp, reps, _p(out), _p(st, _i64p),
                                  ctypes.byref(ms)), "multi_run")
        return out, dict(nodes=int(st[0]), fallbacks=int(st[1]), pairs=int(st[2])), ms.value
