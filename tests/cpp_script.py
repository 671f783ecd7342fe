"""Drives tests/cpp/api_script (the host C++ API) from a test: writes the
parameters, keys and inputs as raw uint64 files, the command script, runs
it, and reads back every stored ciphertext with its metadata."""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "api_script")


class Script:
    def __init__(self, tmp_path, prm, technique=0):
        assert os.path.exists(BIN), "tests/cpp/api_script is not built (run __graft_entry__.build())"
        self.d = str(tmp_path)
        self.prm = prm
        self.lines = []
        self.nfiles = 0
        hdr = [prm.logN, prm.L, prm.K, prm.dnum, 0, 0] + list(prm.q) + list(prm.p)
        np.asarray(hdr, np.uint64).tofile(os.path.join(self.d, "params.bin"))
        self.lines.append("context params.bin %d" % technique)

    def _file(self, arr):
        name = "in%d.bin" % self.nfiles
        self.nfiles += 1
        np.ascontiguousarray(arr, np.uint64).tofile(os.path.join(self.d, name))
        return name

    def rlk(self, key):
        self.lines.append("rlk " + self._file(key))

    def autokey(self, k, key):
        self.lines.append("autokey %d %s" % (k, self._file(key)))

    def ct(self, name, data, deg=1, sf=1.0, slots=0):
        self.lines.append("ct %s %s %d %d %r %d" % (name, self._file(data), data.shape[1], deg, float(sf), slots))

    def pt(self, name, limbs_data, fmt, deg, sf):
        self.lines.append("pt %s %s %d %d %d %r" % (name, self._file(limbs_data), limbs_data.shape[0], fmt, deg,
                                                    float(sf)))

    def encapkeys(self, down, up):
        self.lines.append("encapkeys %s %s" % (self._file(down), self._file(up)))

    def bootsetup_pre(self, pre, cfg):
        """Write BootstrapPlan.export_precomputation() `pre` (config `cfg`)."""
        words = [len(pre["cts"]), len(pre["stc"]), len(pre["cheb"]), cfg.correction, cfg.double_angle,
                 1 if cfg.encaps_h else 0, cfg.slots or 0]
        f64 = lambda x: int(np.array([x], np.float64).view(np.uint64)[0])
        parts = [np.asarray(words + [f64(pre["out_scale"])] + [f64(c) for c in pre["cheb"]], np.uint64)]
        for g in pre["cts"] + pre["stc"]:
            parts.append(np.asarray([g["stride"], g["g"], len(g["babies"])] +
                                    [b % (1 << 64) for b in g["babies"]] + [len(g["terms"])], np.uint64))
            for k, b, level, scale, coef in g["terms"]:
                parts.append(np.asarray([k % (1 << 64), b % (1 << 64), level, f64(scale)], np.uint64))
                parts.append(np.ascontiguousarray(coef, np.uint64).ravel())
        self.lines.append("bootsetup_pre " + self._file(np.concatenate(parts)))

    def op(self, *words):
        self.lines.append(" ".join(repr(w) if isinstance(w, float) else str(w) for w in words))

    def store(self, *names):
        for n in names:
            self.lines.append("store " + n)

    def run(self, timeout=600):
        with open(os.path.join(self.d, "script.txt"), "w") as f:
            f.write("\n".join(self.lines) + "\n")
        print("api_script: %d commands" % len(self.lines), flush=True)
        r = subprocess.run([BIN, self.d], capture_output=True, text=True, timeout=timeout)
        assert r.returncode == 0, r.stderr
        out = {"throws": []}
        N = self.prm.N
        for line in open(os.path.join(self.d, "meta.txt")).read().strip().split("\n"):
            w = line.split()
            if w[0] == "sf":
                out["sf"] = [float(x) for x in w[1:]]
            elif w[0] == "rots":
                out["rots"] = [int(x) for x in w[1:]]
            elif w[0] == "throws":
                out["throws"].append(w[1])
            elif w[0] == "boottime":
                out.setdefault("boottime", []).append((int(w[1]), float(w[2]), float(w[3])))
            else:
                limbs = int(w[1])
                data = np.fromfile(os.path.join(self.d, w[0] + ".bin"), np.uint64).reshape(2, limbs, N)
                out[w[0]] = dict(data=data, limbs=limbs, deg=int(w[2]), sf=float(w[3]), slots=int(w[4]),
                                 ext=int(w[5]))
        return out

    def read_text(self, name):
        return [float(x) for x in open(os.path.join(self.d, name + ".txt")).read().split()]
