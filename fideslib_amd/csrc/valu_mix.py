"""VALU instruction mix of the forward NTT passes, from this source's own
gfx950 assembly (measurement support for the NTT roofline).

Compiles ntt.hip for the device only (`hipcc -S --cuda-device-only`, the
library's own flags) into a temporary directory and counts, per kernel, the
VALU opcodes of its whole body (every `v_` instruction up to the function's
end), with the encoding suffixes _e32/_e64/_dpp/_sdwa dropped.  The passes
are straight-line code (every loop is unrolled), so a static count is the
count each thread executes: 8 butterflies per stage x 8 stages = 64 per
thread for N = 2^16.

Usage: python3 fideslib_amd/csrc/valu_mix.py [> profiles/ntt_fwd_valu_mix.json]
"""
import collections
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-S", "--cuda-device-only"]
# symbol -> (name as rocprofv3 prints it, butterflies per thread)
KERNELS = {
    "_Z13k_ntt_fwd_colILi16ELi0ELi1EEv8LimbSpan9NttTables7NttFuse": ("void k_ntt_fwd_col<16, 0, 1>", 64),
    "_Z13k_ntt_fwd_rowILi16ELi0ELi1EEv8LimbSpan9NttTables7NttFuse": ("void k_ntt_fwd_row<16, 0, 1>", 64),
}


def count(asm_lines, symbol):
    ops, on = collections.Counter(), False
    for ln in asm_lines:
        if ln.startswith(symbol + ":"):
            on = True
            continue
        if not on:
            continue
        if ln.startswith(".Lfunc_end"):
            break
        t = ln.strip().split()
        if t and t[0].startswith("v_"):
            ops[re.sub(r"_(e32|e64|dpp|sdwa)$", "", t[0])] += 1
    if not on:
        raise SystemExit("kernel %s not found in the assembly" % symbol)
    return ops


def main():
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "ntt.s")
        subprocess.run([HIPCC] + FLAGS + [os.path.join(HERE, "ntt.hip"), "-o", out], check=True,
                       stderr=subprocess.DEVNULL)
        lines = open(out).read().split("\n")
    res = {}
    for sym, (name, bfly) in KERNELS.items():
        ops = count(lines, sym)
        res[name] = {"valu_per_thread": sum(ops.values()), "butterflies_per_thread": bfly,
                     "ops": dict(sorted(ops.items(), key=lambda kv: -kv[1]))}
    json.dump({"command": "python3 fideslib_amd/csrc/valu_mix.py > profiles/ntt_fwd_valu_mix.json",
               "source": "fideslib_amd/csrc/ntt.hip", "flags": " ".join(FLAGS), "kernels": res}, sys.stdout, indent=1)
    print()


if __name__ == "__main__":
    main()
