"""ctypes front-end of the CPU oracle (oracle/ckks_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product package.

Arrays are numpy uint64, C-contiguous, shaped [..., N].
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FIDES_ORACLE_SANITIZE=1 selects the AddressSanitizer/UBSan build (make sanitize)
_SANITIZE = os.environ.get("FIDES_ORACLE_SANITIZE") == "1"
_LIB_PATH = os.path.join(_HERE, "build", "libckks_oracle_san.so" if _SANITIZE else "libckks_oracle.so")

u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []) + ["sanitize" if _SANITIZE else "all"],
                       check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c = ctypes.c_uint64
        i = ctypes.c_uint32
        vp = ctypes.c_void_p
        sig = {
            "or_mulmod": (c, [c, c, c]),
            "or_powmod": (c, [c, c, c]),
            "or_invmod": (c, [c, c]),
            "or_brev": (i, [i, i]),
            "or_generator": (c, [c]),
            "or_root_2n": (c, [c, i]),
            "or_psi_tables": (None, [c, c, i, u64p, u64p]),
            "or_ntt": (None, [u64p, i, c, u64p]),
            "or_intt": (None, [u64p, i, c, u64p]),
            "or_ctx_create": (vp, [i, u64p, i, u64p, i, i, u64p]),
            "or_ctx_destroy": (None, [vp]),
            "or_ctx_root": (c, [vp, i]),
            "or_ctx_alpha": (i, [vp]),
            "or_num_digits": (i, [vp, i]),
            "or_ntt_limbs": (None, [vp, u64p, u32p, i]),
            "or_intt_limbs": (None, [vp, u64p, u32p, i]),
            "or_switch_basis": (None, [vp, u64p, u32p, i, u64p, u32p, i]),
            "or_modup_digit": (None, [vp, u64p, i, i, u64p]),
            "or_moddown": (None, [vp, u64p, i, u64p]),
            "or_keyswitch": (None, [vp, u64p, i, u64p, u64p, u64p]),
            "or_tensor": (None, [vp, u64p, u64p, i, u64p]),
            "or_switch_modulus": (c, [c, c, c]),
            "or_rescale_table": (None, [vp, i, u64p]),
            "or_rescale_poly": (None, [vp, u64p, i]),
            "or_rescale_ct": (None, [vp, u64p, i, u64p]),
            "or_mod_raise": (None, [vp, u64p, i, u64p]),
            "or_set_threads": (None, [ctypes.c_int]),
            "or_get_threads": (ctypes.c_int, []),
            "or_automorph": (None, [vp, u64p, u64p, i, i]),
            "or_hmult_relin": (None, [vp, u64p, u64p, i, u64p, u64p]),
            "or_rotate": (None, [vp, u64p, i, u64p, i, u64p]),
            "or_keygen_switch": (None, [vp, u64p, u64p, u64p, u64p, u64p]),
            "or_mul_limbs": (None, [vp, u64p, u64p, u64p, u32p, i]),
            "or_add_limbs": (None, [vp, u64p, u64p, u64p, u32p, i]),
            "or_sub_limbs": (None, [vp, u64p, u64p, u64p, u32p, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(u64p)


def _p32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u32p)


def _u64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


# ---------------------------------------------------------------- scalars
def mulmod(a, b, q):
    return lib().or_mulmod(a, b, q)


def powmod(a, e, q):
    return lib().or_powmod(a, e, q)


def invmod(a, q):
    return lib().or_invmod(a, q)


def root_2n(q: int, logN: int) -> int:
    return lib().or_root_2n(q, logN)


def psi_tables(q: int, psi: int, logN: int):
    N = 1 << logN
    a = np.zeros(N, np.uint64)
    b = np.zeros(N, np.uint64)
    lib().or_psi_tables(q, psi, logN, _p64(a), _p64(b))
    return a, b


def ntt(a: np.ndarray, logN: int, q: int, psi_rev: np.ndarray) -> np.ndarray:
    x = _u64(a).copy()
    lib().or_ntt(_p64(x), logN, q, _p64(_u64(psi_rev)))
    return x


def intt(a: np.ndarray, logN: int, q: int, ipsi_rev: np.ndarray) -> np.ndarray:
    x = _u64(a).copy()
    lib().or_intt(_p64(x), logN, q, _p64(_u64(ipsi_rev)))
    return x


def switch_modulus(a: int, om: int, nm: int) -> int:
    return lib().or_switch_modulus(a, om, nm)


class Context:
    """Oracle context: Q primes q[0..L), special primes p[0..K), dnum digits."""

    def __init__(self, logN: int, q, p, dnum: int, roots=None):
        self.logN = logN
        self.N = 1 << logN
        self.q = [int(x) for x in q]
        self.p = [int(x) for x in p]
        self.L = len(self.q)
        self.K = len(self.p)
        self.dnum = dnum
        self.primes = self.q + self.p
        qa = _u64(self.q)
        pa = _u64(self.p) if self.p else np.zeros(1, np.uint64)
        ra = _u64(roots) if roots is not None else None
        self._h = lib().or_ctx_create(logN, _p64(qa), self.L, _p64(pa), self.K, dnum,
                                      _p64(ra) if ra is not None else None)
        self.alpha = lib().or_ctx_alpha(self._h)
        self.roots = [lib().or_ctx_root(self._h, i) for i in range(self.L + self.K)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_ctx_destroy(h)
            self._h = None

    # -- helpers
    def num_digits(self, l: int) -> int:
        return lib().or_num_digits(self._h, l)

    def pids_q(self, l: int) -> np.ndarray:
        return np.arange(l, dtype=np.uint32)

    def pids_qp(self, l: int) -> np.ndarray:
        return np.concatenate([np.arange(l), self.L + np.arange(self.K)]).astype(np.uint32)

    # -- limb transforms
    def ntt_limbs(self, data: np.ndarray, pids) -> np.ndarray:
        x = _u64(data).copy()
        pid = np.ascontiguousarray(np.asarray(pids, np.uint32))
        assert x.size == len(pid) * self.N
        lib().or_ntt_limbs(self._h, _p64(x), _p32(pid), len(pid))
        return x

    def intt_limbs(self, data: np.ndarray, pids) -> np.ndarray:
        x = _u64(data).copy()
        pid = np.ascontiguousarray(np.asarray(pids, np.uint32))
        assert x.size == len(pid) * self.N
        lib().or_intt_limbs(self._h, _p64(x), _p32(pid), len(pid))
        return x

    def switch_basis(self, coef: np.ndarray, src, dst) -> np.ndarray:
        s = np.ascontiguousarray(np.asarray(src, np.uint32))
        d = np.ascontiguousarray(np.asarray(dst, np.uint32))
        x = _u64(coef)
        out = np.zeros((len(d), self.N), np.uint64)
        lib().or_switch_basis(self._h, _p64(x), _p32(s), len(s), _p64(out), _p32(d), len(d))
        return out

    # -- key switching pieces
    def modup_digit(self, d: np.ndarray, l: int, j: int) -> np.ndarray:
        x = _u64(d)
        assert x.size == l * self.N
        out = np.zeros((l + self.K, self.N), np.uint64)
        lib().or_modup_digit(self._h, _p64(x), l, j, _p64(out))
        return out

    def moddown(self, ext: np.ndarray, l: int) -> np.ndarray:
        x = _u64(ext)
        assert x.size == (l + self.K) * self.N
        out = np.zeros((l, self.N), np.uint64)
        lib().or_moddown(self._h, _p64(x), l, _p64(out))
        return out

    def keyswitch(self, d: np.ndarray, l: int, key: np.ndarray):
        x = _u64(d)
        k = _u64(key)
        assert k.size == 2 * self.dnum * (self.L + self.K) * self.N
        o0 = np.zeros((l, self.N), np.uint64)
        o1 = np.zeros((l, self.N), np.uint64)
        lib().or_keyswitch(self._h, _p64(x), l, _p64(k), _p64(o0), _p64(o1))
        return o0, o1

    def tensor(self, a: np.ndarray, b: np.ndarray, l: int) -> np.ndarray:
        out = np.zeros((3, l, self.N), np.uint64)
        lib().or_tensor(self._h, _p64(_u64(a)), _p64(_u64(b)), l, _p64(out))
        return out

    def rescale_table(self, l: int):
        out = np.zeros(max(l - 1, 1), np.uint64)
        lib().or_rescale_table(self._h, l, _p64(out))
        return [int(v) for v in out[: l - 1]]

    def rescale_ct(self, ct: np.ndarray, l: int) -> np.ndarray:
        out = np.zeros((2, l - 1, self.N), np.uint64)
        lib().or_rescale_ct(self._h, _p64(_u64(ct)), l, _p64(out))
        return out

    def automorph(self, x: np.ndarray, k: int) -> np.ndarray:
        a = _u64(x)
        n = a.size // self.N
        out = np.zeros_like(a)
        lib().or_automorph(self._h, _p64(a), _p64(out), n, k)
        return out

    def hmult_relin(self, a: np.ndarray, b: np.ndarray, l: int, rlk: np.ndarray) -> np.ndarray:
        out = np.zeros((2, l, self.N), np.uint64)
        lib().or_hmult_relin(self._h, _p64(_u64(a)), _p64(_u64(b)), l, _p64(_u64(rlk)), _p64(out))
        return out

    def rotate(self, ct: np.ndarray, l: int, key: np.ndarray, k: int) -> np.ndarray:
        out = np.zeros((2, l, self.N), np.uint64)
        lib().or_rotate(self._h, _p64(_u64(ct)), l, _p64(_u64(key)), k, _p64(out))
        return out

    def keygen_switch(self, s_new: np.ndarray, s_old: np.ndarray, a: np.ndarray, e: np.ndarray) -> np.ndarray:
        LK = self.L + self.K
        key = np.zeros((2, self.dnum, LK, self.N), np.uint64)
        lib().or_keygen_switch(self._h, _p64(_u64(s_new)), _p64(_u64(s_old)), _p64(_u64(a)), _p64(_u64(e)),
                               _p64(key))
        return key

    def mod_raise(self, ct: np.ndarray, l_out: int) -> np.ndarray:
        """[2][1][N] at q_0 -> [2][l_out][N] (centred lift of each coefficient)."""
        out = np.zeros((2, l_out, self.N), np.uint64)
        lib().or_mod_raise(self._h, _p64(_u64(ct)), l_out, _p64(out))
        return out

    def mul_limbs(self, x, y, pids):
        pid = np.ascontiguousarray(np.asarray(pids, np.uint32))
        out = np.zeros((len(pid), self.N), np.uint64)
        lib().or_mul_limbs(self._h, _p64(_u64(x)), _p64(_u64(y)), _p64(out), _p32(pid), len(pid))
        return out

    def add_limbs(self, x, y, pids):
        pid = np.ascontiguousarray(np.asarray(pids, np.uint32))
        out = np.zeros((len(pid), self.N), np.uint64)
        lib().or_add_limbs(self._h, _p64(_u64(x)), _p64(_u64(y)), _p64(out), _p32(pid), len(pid))
        return out

    def sub_limbs(self, x, y, pids):
        pid = np.ascontiguousarray(np.asarray(pids, np.uint32))
        out = np.zeros((len(pid), self.N), np.uint64)
        lib().or_sub_limbs(self._h, _p64(_u64(x)), _p64(_u64(y)), _p64(out), _p32(pid), len(pid))
        return out


def set_threads(n: int) -> int:
    """Set the oracle's OpenMP worker count; returns the previous one."""
    prev = lib().or_get_threads()
    lib().or_set_threads(int(n))
    return prev
