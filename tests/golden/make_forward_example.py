"""Writes tests/golden/forward_example.npz from the reference's own example
inputs: examples/forward (config.ini, sources.dat, stations.dat, the
model_true vp/vs/rho/x/z files and the model_init vp/vs/rho files, int32
npt header + float32) and the config.ini files of examples/adjoint and
examples/inverse-checker, which inherit from it.  Data only: the file
holds the arrays and the text inputs verbatim.

    python tests/golden/make_forward_example.py [/root/reference]
"""
import os
import sys

import numpy as np

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
EX = os.path.join(ROOT, "examples", "forward")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "forward_example.npz")


def read_bin(name, model="model_true"):
    a = np.fromfile(os.path.join(EX, model, "proc000000_%s.bin" % name), np.float32)
    npt = int(a[:1].view(np.int32)[0])
    assert a.size == npt + 1
    return a[1:]


def main():
    arrays = {k: read_bin(k) for k in ("vp", "vs", "rho", "x", "z")}
    for k in ("vp", "vs", "rho"):
        arrays["init_" + k] = read_bin(k, "model_init")
    for k in ("config.ini", "sources.dat", "stations.dat"):
        arrays[k.replace(".", "_")] = np.array(open(os.path.join(EX, k)).read())
    for ex in ("adjoint", "inverse-checker"):
        cfg = open(os.path.join(ROOT, "examples", ex, "config.ini")).read()
        arrays[ex.replace("-", "_") + "_config_ini"] = np.array(cfg)
    np.savez_compressed(OUT, **arrays)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
