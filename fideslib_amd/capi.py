"""ctypes binding of the C ABI in include/fideslib_amd.h.

Loads the in-tree libfideslib_amd.so (built by fideslib_amd/build.py).  There
is no fallback: if the library is missing or cannot be loaded, loading
raises, so no operation can silently run elsewhere.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfideslib_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "fideslib_amd.h")

FIDES_OK = 0


class FidesError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("logN", ctypes.c_uint32),
        ("L", ctypes.c_uint32),
        ("K", ctypes.c_uint32),
        ("dnum", ctypes.c_uint32),
        ("q", ctypes.POINTER(ctypes.c_uint64)),
        ("p", ctypes.POINTER(ctypes.c_uint64)),
        ("psi", ctypes.POINTER(ctypes.c_uint64)),
    ]


class Tuning(ctypes.Structure):
    """Launch-shape settings of a context (fides_tuning); none changes a result."""
    FIELDS = ("conv_chunk", "group_by_prime", "fuse_row_dot")  # each per size class 0..2
    _fields_ = [(f, ctypes.c_uint32 * 3) for f in FIELDS]

    def as_dict(self) -> dict:
        return {f: [int(v) for v in getattr(self, f)] for f in self.FIELDS}

    @classmethod
    def from_dict(cls, d: dict) -> "Tuning":
        t = cls()
        for f in cls.FIELDS:
            if len(d[f]) != 3:
                raise FidesError("tuning field %s needs 3 entries (one per size class)" % f)
            for i, v in enumerate(d[f]):
                getattr(t, f)[i] = int(v)
        return t


_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_int = ctypes.c_int

# name -> (restype, argtypes); the list is also the export contract the CPU
# tests check against the header.
SIGNATURES = {
    "fides_last_error": (ctypes.c_char_p, []),
    "fides_context_create": (_int, [ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
    "fides_context_destroy": (None, [_vp]),
    "fides_context_info": (_int, [_vp, _u32p, _u32p, _u32p, _u32p, _u32p]),
    "fides_context_roots": (_int, [_vp, _u64p]),
    "fides_device_alloc": (_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "fides_device_free": (_int, [_vp]),
    "fides_copy_to_device": (_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "fides_copy_to_host": (_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "fides_copy_device": (_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "fides_stream_sync": (_int, [_vp]),
    "fides_stream_create": (_int, [ctypes.POINTER(_vp)]),
    "fides_stream_destroy": (_int, [_vp]),
    "fides_graph_begin": (_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_size_t]),
    "fides_graph_begin_shared": (_int, [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_size_t]),
    "fides_graph_end": (_int, [_vp, ctypes.POINTER(_vp)]),
    "fides_graph_launch": (_int, [_vp, _vp]),
    "fides_graph_destroy": (_int, [_vp]),
    "fides_tuning_get": (_int, [_vp, ctypes.POINTER(Tuning)]),
    "fides_tuning_set": (_int, [_vp, ctypes.POINTER(Tuning)]),
    "fides_autotune": (_int, [_vp, _u32, _u32p, _u32, ctypes.POINTER(Tuning)]),
    "fides_ntt": (_int, [_vp, _vp, _u32p, _u32, _vp]),
    "fides_intt": (_int, [_vp, _vp, _u32p, _u32, _vp]),
    "fides_modup": (_int, [_vp, _vp, _u32, _u32, _vp, _vp]),
    "fides_moddown": (_int, [_vp, _vp, _u32, _vp, _vp]),
    "fides_key_upload": (_int, [_vp, _u64p, ctypes.POINTER(_vp)]),
    "fides_key_destroy": (None, [_vp]),
    "fides_keyswitch": (_int, [_vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    "fides_keyswitch_noalloc": (_int, [_vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    "fides_hmult": (_int, [_vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "fides_square": (_int, [_vp, _vp, _u32, _vp, _vp, _vp]),
    "fides_rescale": (_int, [_vp, _vp, _u32, _vp, _vp]),
    "fides_rotate": (_int, [_vp, _vp, _u32, _vp, _u32, _vp, _vp]),
    "fides_automorph": (_int, [_vp, _vp, _u32, _u32, _vp, _vp]),
    "fides_hmult_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u32p, _vp, _int,
                                 ctypes.POINTER(_vp), _vp]),
    "fides_rotate_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_vp), _u32p,
                                  ctypes.POINTER(_vp), _vp]),
    "fides_rotate_hoisted": (_int, [_vp, _vp, _u32, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_plaintext_prepare": (_int, [_vp, _vp, _u32, _vp]),
    "fides_ptdot_batch": (_int, [_vp, _u32, _u32p, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u32p,
                                 ctypes.POINTER(_vp), _vp]),
    "fides_mod_raise_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_key_download": (_int, [_vp, _u64p]),
    "fides_rescale_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_rotate_hoisted_ext": (_int, [_vp, _vp, _u32, _u32, ctypes.POINTER(_vp), _u32p, _int,
                                        ctypes.POINTER(_vp), _vp]),
    "fides_ext_raise_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_ext_add_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u32p, _int,
                                   ctypes.POINTER(_vp), _vp]),
    "fides_ext_moddown_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_plaintext_embed": (_int, [_vp, _vp, _u32, _vp, _vp]),
    "fides_lin_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_u64p),
                               ctypes.POINTER(_u64p), _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_lincomb_batch": (_int, [_vp, _u32, _u32p, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_u64p),
                                   ctypes.POINTER(_u64p), _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_drop_limbs_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, _u32p, ctypes.POINTER(_vp), _vp]),
    "fides_add_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u32p, _int,
                               ctypes.POINTER(_vp), _vp]),
    "fides_mul_const_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_u64p),
                                     ctypes.POINTER(_vp), _vp]),
    "fides_add_const_batch": (_int, [_vp, _u32, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(_u64p),
                                     ctypes.POINTER(_vp), _vp]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FidesError("libfideslib_amd.so is not built (run fideslib_amd/build.py): " + LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != FIDES_OK:
        msg = load().fides_last_error()
        raise FidesError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


# ------------------------------------------------------------------ device buffers
class DeviceBuffer:
    """A device allocation of uint64 words owned by this object."""

    def __init__(self, nwords: int):
        self.nwords = int(nwords)
        p = _vp()
        check(load().fides_device_alloc(max(1, self.nwords) * 8, ctypes.byref(p)), "fides_device_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = cls(a.size)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a, dtype=np.uint64)
        assert a.size <= self.nwords
        check(load().fides_copy_to_device(self.ptr, a.ctypes.data, a.size * 8, _sptr(stream)), "copy_to_device")

    def download(self, shape=None, stream=None, nwords=None) -> np.ndarray:
        n = self.nwords if nwords is None else nwords
        out = np.empty(n, np.uint64)
        check(load().fides_copy_to_host(out.ctypes.data, self.ptr, n * 8, _sptr(stream)), "copy_to_host")
        return out.reshape(shape) if shape is not None else out

    def at(self, word_offset: int) -> int:
        return self.ptr + 8 * int(word_offset)

    def free(self) -> None:
        if getattr(self, "ptr", None):
            load().fides_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ context
class Context:
    """A device context for one parameter set (Q primes, P primes, dnum)."""

    def __init__(self, logN: int, q, p, dnum: int, psi=None):
        lib = load()
        self.logN = logN
        self.N = 1 << logN
        self.q = [int(x) for x in q]
        self.p = [int(x) for x in p]
        self.L, self.K, self.dnum = len(self.q), len(self.p), dnum
        self._qa = np.ascontiguousarray(self.q, np.uint64)
        self._pa = np.ascontiguousarray(self.p, np.uint64)
        self._psi = None if psi is None else np.ascontiguousarray([int(x) for x in psi], np.uint64)
        prm = Params(logN, self.L, self.K, dnum, self._qa.ctypes.data_as(_u64p), self._pa.ctypes.data_as(_u64p),
                     self._psi.ctypes.data_as(_u64p) if self._psi is not None else None)
        h = _vp()
        check(lib.fides_context_create(ctypes.byref(prm), ctypes.byref(h)), "fides_context_create")
        self.h = h.value
        a = _u32()
        lib.fides_context_info(self.h, None, None, None, None, ctypes.byref(a))
        self.alpha = a.value
        roots = np.zeros(self.L + self.K, np.uint64)
        check(lib.fides_context_roots(self.h, roots.ctypes.data_as(_u64p)), "roots")
        self.roots = [int(r) for r in roots]

    def close(self):
        if getattr(self, "h", None):
            load().fides_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def digits(self, l: int) -> int:
        return (l + self.alpha - 1) // self.alpha

    # ---- raw limb ops
    def ntt(self, buf: DeviceBuffer, pids, stream=None, inverse=False, offset=0):
        pid = np.ascontiguousarray(pids, np.uint32)
        f = load().fides_intt if inverse else load().fides_ntt
        check(f(self.h, buf.at(offset), pid.ctypes.data_as(_u32p), len(pid), _sptr(stream)), "fides_ntt")

    def intt(self, buf: DeviceBuffer, pids, stream=None, offset=0):
        self.ntt(buf, pids, stream, inverse=True, offset=offset)

    def modup(self, inp: DeviceBuffer, l: int, digit: int, out: DeviceBuffer, stream=None):
        check(load().fides_modup(self.h, inp.ptr, l, digit, out.ptr, _sptr(stream)), "fides_modup")

    def moddown(self, inp: DeviceBuffer, l: int, out: DeviceBuffer, stream=None):
        check(load().fides_moddown(self.h, inp.ptr, l, out.ptr, _sptr(stream)), "fides_moddown")

    # ---- keys
    def upload_key(self, key: np.ndarray) -> "Key":
        return Key(self, key)

    def keyswitch(self, inp, l, key, out0, out1, stream=None):
        check(load().fides_keyswitch(self.h, inp.ptr, l, key.h, out0.ptr, out1.ptr, _sptr(stream)), "fides_keyswitch")

    def keyswitch_noalloc(self, inp, l, key, out0, out1, stream=None):
        """fides_keyswitch with the library's static device scratch (no allocation)."""
        check(load().fides_keyswitch_noalloc(self.h, inp.ptr, l, key.h, out0.ptr, out1.ptr, _sptr(stream)),
              "fides_keyswitch_noalloc")

    # ---- ciphertext ops
    def hmult(self, a, b, l, rlk, out, stream=None):
        check(load().fides_hmult(self.h, _ptr(a), _ptr(b), l, rlk.h, _ptr(out), _sptr(stream)), "fides_hmult")

    def square(self, a, l, rlk, out, stream=None):
        check(load().fides_square(self.h, _ptr(a), l, rlk.h, _ptr(out), _sptr(stream)), "fides_square")

    def rescale(self, inp, l, out, stream=None):
        check(load().fides_rescale(self.h, _ptr(inp), l, _ptr(out), _sptr(stream)), "fides_rescale")

    def rotate(self, inp, l, key, k, out, stream=None):
        check(load().fides_rotate(self.h, _ptr(inp), l, key.h, k, _ptr(out), _sptr(stream)), "fides_rotate")

    def automorph(self, inp, nlimbs, k, out, stream=None):
        check(load().fides_automorph(self.h, _ptr(inp), nlimbs, k, _ptr(out), _sptr(stream)), "fides_automorph")

    def hmult_batch(self, a, b, levels, rlk, out, rescale=True, stream=None):
        n = len(levels)
        A = (_vp * n)(*[_ptr(x) for x in a])
        B = (_vp * n)(*[_ptr(x) for x in b])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        check(load().fides_hmult_batch(self.h, n, A, B, lv.ctypes.data_as(_u32p), rlk.h, int(bool(rescale)), O,
                                       _sptr(stream)), "fides_hmult_batch")

    def rotate_batch(self, inp, levels, keys, kidx, out, stream=None):
        n = len(levels)
        I = (_vp * n)(*[_ptr(x) for x in inp])
        O = (_vp * n)(*[_ptr(x) for x in out])
        Kh = (_vp * n)(*[k.h for k in keys])
        lv = np.ascontiguousarray(levels, np.uint32)
        kx = np.ascontiguousarray(kidx, np.uint32)
        check(load().fides_rotate_batch(self.h, n, I, lv.ctypes.data_as(_u32p), Kh, kx.ctypes.data_as(_u32p), O,
                                        _sptr(stream)), "fides_rotate_batch")

    def rotate_hoisted(self, inp, level, keys, kidx, out, stream=None):
        """n rotations of one ciphertext sharing one ModUp; keys[i] may be None when kidx[i] == 1."""
        n = len(kidx)
        K = (_vp * n)(*[None if k is None else k.h for k in keys])
        O = (_vp * n)(*[_ptr(x) for x in out])
        ki = np.ascontiguousarray(kidx, np.uint32)
        check(load().fides_rotate_hoisted(self.h, _ptr(inp), level, n, K, ki.ctypes.data_as(_u32p), O, _sptr(stream)),
              "fides_rotate_hoisted")

    def plaintext_prepare(self, pt, level, stream=None):
        check(load().fides_plaintext_prepare(self.h, _ptr(pt), level, _sptr(stream)), "fides_plaintext_prepare")

    def ptdot_batch(self, terms, levels, out, stream=None):
        """out[i] = sum_j ct_ij (*) pt_ij for terms[i] = [(ct, pt), ...] (prepared plaintexts)."""
        n = len(levels)
        cnt = np.ascontiguousarray([len(t) for t in terms], np.uint32)
        flat = [x for t in terms for x in t]
        C = (_vp * len(flat))(*[_ptr(c) for c, _ in flat])
        P = (_vp * len(flat))(*[_ptr(p) for _, p in flat])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        check(load().fides_ptdot_batch(self.h, n, cnt.ctypes.data_as(_u32p), C, P, lv.ctypes.data_as(_u32p), O,
                                       _sptr(stream)), "fides_ptdot_batch")

    # -- extended basis Q_l u P: [2][l+K][N] buffers
    def rotate_hoisted_ext(self, inp, level, keys, kidx, out, add_first=True, stream=None):
        n = len(kidx)
        K = (_vp * n)(*[None if k is None else k.h for k in keys])
        O = (_vp * n)(*[_ptr(x) for x in out])
        ki = np.ascontiguousarray(kidx, np.uint32)
        check(load().fides_rotate_hoisted_ext(self.h, _ptr(inp), level, n, K, ki.ctypes.data_as(_u32p),
                                              int(bool(add_first)), O, _sptr(stream)), "fides_rotate_hoisted_ext")

    def _ext_unary(self, fn, inp, levels, out, stream, what):
        n = len(levels)
        I = (_vp * n)(*[_ptr(x) for x in inp])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        check(fn(self.h, n, I, lv.ctypes.data_as(_u32p), O, _sptr(stream)), what)

    def ext_raise_batch(self, inp, levels, out, stream=None):
        self._ext_unary(load().fides_ext_raise_batch, inp, levels, out, stream, "fides_ext_raise_batch")

    def ext_moddown_batch(self, inp, levels, out, stream=None):
        self._ext_unary(load().fides_ext_moddown_batch, inp, levels, out, stream, "fides_ext_moddown_batch")

    def ext_add_batch(self, a, b, levels, out, subtract=False, stream=None):
        n = len(levels)
        A = (_vp * n)(*[_ptr(x) for x in a])
        B = (_vp * n)(*[_ptr(x) for x in b])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        check(load().fides_ext_add_batch(self.h, n, A, B, lv.ctypes.data_as(_u32p), int(bool(subtract)), O,
                                         _sptr(stream)), "fides_ext_add_batch")

    def plaintext_embed(self, pt, level, out, stream=None):
        check(load().fides_plaintext_embed(self.h, _ptr(pt), level, _ptr(out), _sptr(stream)), "fides_plaintext_embed")

    def mod_raise_batch(self, inp, l_out, out, stream=None):
        n = len(l_out)
        I = (_vp * n)(*[_ptr(x) for x in inp])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lo = np.ascontiguousarray(l_out, np.uint32)
        check(load().fides_mod_raise_batch(self.h, n, I, lo.ctypes.data_as(_u32p), O, _sptr(stream)),
              "fides_mod_raise_batch")

    def rescale_batch(self, inp, levels, out, stream=None):
        n = len(levels)
        I = (_vp * n)(*[_ptr(x) for x in inp])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        check(load().fides_rescale_batch(self.h, n, I, lv.ctypes.data_as(_u32p), O, _sptr(stream)), "fides_rescale_batch")

    def add_batch(self, a, b, levels, out, subtract=False, stream=None):
        n = len(levels)
        A = (_vp * n)(*[_ptr(x) for x in a])
        B = (_vp * n)(*[_ptr(x) for x in b])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        check(load().fides_add_batch(self.h, n, A, B, lv.ctypes.data_as(_u32p), int(bool(subtract)), O, _sptr(stream)),
              "fides_add_batch")

    def lin_batch(self, a, b, levels, mul, add, sub, out, stream=None):
        """out[i] = a[i] * mul[i] (+/-) b[i] + add[i]; b[i], mul[i], add[i] may be None.
        mul/add are Python integers (any sign or size)."""
        n = len(levels)
        A = (_vp * n)(*[_ptr(x) for x in a])
        B = (_vp * n)(*[None if x is None else _ptr(x) for x in b])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        keep = []

        def res(vals):
            arr = (_u64p * n)()
            for i, (c, l) in enumerate(zip(vals, levels)):
                if c is not None:
                    r = np.ascontiguousarray([int(c) % self.q[j] for j in range(min(l, self.L))], np.uint64)
                    keep.append(r)
                    arr[i] = r.ctypes.data_as(_u64p)
            return arr

        M, C = res(mul), res(add)
        S = np.ascontiguousarray([1 if x else 0 for x in sub], np.uint32)
        check(load().fides_lin_batch(self.h, n, A, B, lv.ctypes.data_as(_u32p), M, C, S.ctypes.data_as(_u32p), O,
                                     _sptr(stream)), "fides_lin_batch")

    def lincomb_batch(self, terms, levels, add, out, stream=None):
        """out[i] = sum_j m_ij * x_ij + add[i]; terms[i] = [(x, x_level, m), ...] with Python-int
        constants; add[i] a Python int or None."""
        n = len(levels)
        cnt = np.ascontiguousarray([len(t) for t in terms], np.uint32)
        flat = [(x, xl, m, l) for t, l in zip(terms, levels) for x, xl, m in t]
        X = (_vp * len(flat))(*[_ptr(x) for x, _, _, _ in flat])
        XL = np.ascontiguousarray([xl for _, xl, _, _ in flat], np.uint32)
        keep = []
        M = (_u64p * len(flat))()
        for k, (_, _, m, l) in enumerate(flat):
            r = np.ascontiguousarray([int(m) % self.q[j] for j in range(min(l, self.L))], np.uint64)
            keep.append(r)
            M[k] = r.ctypes.data_as(_u64p)
        A = (_u64p * n)()
        for i, (a, l) in enumerate(zip(add, levels)):
            if a is not None:
                r = np.ascontiguousarray([int(a) % self.q[j] for j in range(min(l, self.L))], np.uint64)
                keep.append(r)
                A[i] = r.ctypes.data_as(_u64p)
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        check(load().fides_lincomb_batch(self.h, n, cnt.ctypes.data_as(_u32p), X, XL.ctypes.data_as(_u32p), M, A,
                                         lv.ctypes.data_as(_u32p), O, _sptr(stream)), "fides_lincomb_batch")

    def drop_limbs_batch(self, inp, l_in, l_out, out, stream=None):
        n = len(l_in)
        I = (_vp * n)(*[_ptr(x) for x in inp])
        O = (_vp * n)(*[_ptr(x) for x in out])
        li = np.ascontiguousarray(l_in, np.uint32)
        lo = np.ascontiguousarray(l_out, np.uint32)
        check(load().fides_drop_limbs_batch(self.h, n, I, li.ctypes.data_as(_u32p), lo.ctypes.data_as(_u32p), O,
                                            _sptr(stream)), "fides_drop_limbs_batch")

    def _const_call(self, fn, a, levels, consts, out, stream, what):
        n = len(levels)
        A = (_vp * n)(*[_ptr(x) for x in a])
        O = (_vp * n)(*[_ptr(x) for x in out])
        lv = np.ascontiguousarray(levels, np.uint32)
        res = [np.ascontiguousarray([int(c) % self.q[j] for j in range(min(l, self.L))], np.uint64)
               for c, l in zip(consts, levels)]
        R = (_u64p * n)(*[r.ctypes.data_as(_u64p) for r in res])
        check(fn(self.h, n, A, lv.ctypes.data_as(_u32p), R, O, _sptr(stream)), what)

    def mul_const_batch(self, a, levels, consts, out, stream=None):
        """out[i] = a[i] * consts[i] for Python integers (any sign or size)."""
        self._const_call(load().fides_mul_const_batch, a, levels, consts, out, stream, "fides_mul_const_batch")

    def add_const_batch(self, a, levels, consts, out, stream=None):
        """out[i] = a[i] + consts[i] (constant polynomial) for Python integers."""
        self._const_call(load().fides_add_const_batch, a, levels, consts, out, stream, "fides_add_const_batch")

    def copy(self, dst, src, nwords, stream=None):
        check(load().fides_copy_device(_ptr(dst), _ptr(src), int(nwords) * 8, _sptr(stream)), "fides_copy_device")

    def sync(self, stream=None):
        check(load().fides_stream_sync(_sptr(stream)), "sync")

    # -- graph capture: operations issued on `stream` until graph_end are
    #    recorded, not run
    def graph_begin(self, stream: "Stream", scratch_bytes: int, task_bytes: int, scratch: int | None = None) -> None:
        """Start capturing operations issued on `stream`; `scratch` (a device
        pointer) makes the graph use the caller's key-switch scratch of
        scratch_bytes instead of allocating its own."""
        if scratch is None:
            check(load().fides_graph_begin(self.h, stream.h, int(scratch_bytes), int(task_bytes)),
                  "fides_graph_begin")
        else:
            check(load().fides_graph_begin_shared(self.h, stream.h, scratch, int(scratch_bytes), int(task_bytes)),
                  "fides_graph_begin_shared")

    def graph_end(self) -> "Graph":
        h = _vp()
        check(load().fides_graph_end(self.h, ctypes.byref(h)), "fides_graph_end")
        return Graph(h.value)

    # -- tuning (launch shapes only; results are the same under every setting)
    @property
    def tuning(self) -> dict:
        t = Tuning()
        check(load().fides_tuning_get(self.h, ctypes.byref(t)), "fides_tuning_get")
        return t.as_dict()

    @tuning.setter
    def tuning(self, d: dict) -> None:
        t = Tuning.from_dict(d)
        check(load().fides_tuning_set(self.h, ctypes.byref(t)), "fides_tuning_set")

    def autotune(self, levels, reps: int = 3) -> dict:
        """Time candidate settings on a synthetic batch of HMult+Rescale and
        Rotate at these limb counts; keep and return the fastest."""
        lv = (ctypes.c_uint32 * len(levels))(*[int(x) for x in levels])
        t = Tuning()
        check(load().fides_autotune(self.h, len(levels), lv, int(reps), ctypes.byref(t)), "fides_autotune")
        return t.as_dict()


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        h = _vp()
        check(load().fides_stream_create(ctypes.byref(h)), "fides_stream_create")
        self.h = h.value

    def sync(self):
        check(load().fides_stream_sync(self.h), "sync")

    def close(self):
        if getattr(self, "h", None):
            load().fides_stream_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured sequence of operations (see Context.graph_begin)."""

    def __init__(self, h):
        self.h = h

    def launch(self, stream: "Stream") -> None:
        check(load().fides_graph_launch(self.h, stream.h), "fides_graph_launch")

    def close(self):
        if getattr(self, "h", None):
            load().fides_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sptr(stream):
    return stream.h if isinstance(stream, Stream) else stream


def _ptr(x) -> int:
    return x.ptr if isinstance(x, DeviceBuffer) else int(x)


class Key:
    def __init__(self, ctx: Context, key: np.ndarray):
        k = np.ascontiguousarray(key, np.uint64)
        assert k.size == 2 * ctx.dnum * (ctx.L + ctx.K) * ctx.N, "key must be [2][dnum][L+K][N]"
        h = _vp()
        check(load().fides_key_upload(ctx.h, k.ctypes.data_as(_u64p), ctypes.byref(h)), "fides_key_upload")
        self.h = h.value
        self.ctx = ctx

    def close(self):
        if getattr(self, "h", None):
            load().fides_key_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
