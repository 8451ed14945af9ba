"""Writes the per-GPU launch-shape table of the Coulomb kernel.

The reference ships per-GPU launch tables read at start-up
(libraries/NBODY/<GPU>_NB0DY.dat, read by bin/nbody.cpp:275-321).  This is
their counterpart: sbk_autotune times every launch shape on the benchmark
workloads, and the lines ("device ions clouds precision targets/thread
source-groups slices local-coords kernel ms") are written to
simbuca_amd/tuning/<arch>.tsv.
The engine applies the fastest line that matches its device, ion count,
cloud count and precision with sbk_load_tuning.

Run on the GPU:  python simbuca_amd/tuning/make_table.py [out.tsv]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import simbuca_amd as S  # noqa: E402
from simbuca_amd import synth  # noqa: E402


def main():
    arch = S.device_info()["name"].split(":")[0]
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, f"{arch}.tsv")
    if os.path.exists(out):
        os.remove(out)
    for k in (1, 2, 3, 4, 5):
        w = synth.workload(k)
        for precision in sorted({"fp64", "fp64_rsq",
                                 "mixed_kahan" if w.precision == "mixed_kahan" else "mixed"}):
            e = S.Engine(precision=precision)
            for c in w.clouds:
                e.add_cloud(c.pos, c.vel, c.mass, c.charge, c.fields)
            shape = e.autotune(precision, out)
            e.close()
            print(f"workload {k} {precision}: {shape}", flush=True)


if __name__ == "__main__":
    main()
