"""Writes tests/golden/reference_probe.json from the reference probe.

The probe (oracle/ref_probe/probe.cpp) is compiled by `make -C oracle ref`
from the reference's own SPHKernels.cpp, GaussQuadrature.cpp and vendored
Eigen under /root/reference; its JSON output holds the golden vectors:
kernel table samples and kernel values, the p=30 Gauss-Legendre rule,
GaussQuadrature::integrate results, the volume-map integrand of a half-space
wall and Eigen CG with the block-Jacobi preconditioner on a fixed system.
Run it in a container that has /root/reference; the fixture is committed.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_probe")], check=True,
                         capture_output=True, text=True).stdout
    d = json.loads(out)
    cg = d["cg"]
    shared = {k: cg[0][k] for k in ("A", "b", "x0", "n", "maxIter")}
    d["cg"] = {"system": shared,
               "runs": [{k: c[k] for k in ("tol", "iterations", "error", "x")} for c in cg]}
    path = os.path.join(HERE, "reference_probe.json")
    with open(path, "w") as f:
        json.dump(d, f, separators=(",", ":"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    sys.exit(main())
