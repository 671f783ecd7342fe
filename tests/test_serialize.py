"""Binary serialization of the host API objects (include/fideslib/serialize.hpp).

CPU: RawParams, RawCipherText and RawKeySwitchKey round-trip bit for bit
through files; a file of another type, a truncated file and the JSON format
are refused (tests/cpp/api_check.cpp, mode serial-host).
GPU: a whole context (parameters + relinearization and rotation keys) and two
ciphertexts go through files; the deserialized context computes
Rotate(Rescale(EvalMult)) with the same bits as the original (mode serial),
which also equal the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from fideslib_amd import params
from oracle import ckks_codec as C
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "api_check")


def _write_inputs(d, seed):
    prm = params.config1()
    N, l, rot = prm.N, prm.L, 3
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, seed)
    rng = np.random.default_rng(seed)
    z = [rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2) for _ in range(2)]
    cts = [keys.encrypt(C.encode(v, 2.0 ** prm.scale_bits, N), l) for v in z]
    rlk = keys.relin_key()
    rk = keys.rotation_key(pow(5, rot, 2 * N))
    np.asarray([prm.logN, prm.L, prm.K, prm.dnum, l, rot] + list(prm.q) + list(prm.p), np.uint64).tofile(
        os.path.join(d, "params.bin"))
    rlk.astype(np.uint64).tofile(os.path.join(d, "rlk.bin"))
    rk.astype(np.uint64).tofile(os.path.join(d, "rot.bin"))
    cts[0].astype(np.uint64).tofile(os.path.join(d, "c1.bin"))
    cts[1].astype(np.uint64).tofile(os.path.join(d, "c2.bin"))
    return prm, octx, cts, rlk, rk, rot


def test_serialize_host_structs(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/api_check is not built (run __graft_entry__.build())"
    _write_inputs(str(tmp_path), 3)
    r = subprocess.run([BIN, str(tmp_path), "serial-host"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "serial-host: ok" in r.stdout


@pytest.mark.gpu
def test_serialize_context_roundtrip(tmp_path):
    prm, octx, cts, rlk, rk, rot = _write_inputs(str(tmp_path), 4)
    r = subprocess.run([BIN, str(tmp_path), "serial"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    l = prm.L
    got = np.fromfile(os.path.join(str(tmp_path), "serial_result.bin"), np.uint64).reshape(2, l - 1, prm.N)
    want = octx.rotate(octx.rescale_ct(octx.hmult_relin(cts[0], cts[1], l, rlk), l), l - 1, rk, pow(5, rot, 2 * prm.N))
    assert np.array_equal(got, want)
