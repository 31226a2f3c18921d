"""bench.py's command line on the CPU: the one-GPU refusal, and the models of
each single-model workload (the 16^3 regular grid, the envelope, the stress
envelope) at small sizes."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_workload_models():
    sys.path.insert(0, ROOT)
    import bench
    reg = bench.build_workload(4, "regular")
    assert reg.ncells == 64 and reg.nrays == 12 and reg.nfreqs == 39 and reg.nlev == 2
    assert not np.any(reg.velocity)
    env = bench.build_workload(4, "envelope")
    st = bench.build_workload(4, "stress")
    for m in (env, st):
        assert m.ncells == 64 and m.nrays == 48 and m.nfreqs == 186 and m.nlev == 7
    # the stress envelope: n0 1e17 against 1e14 m^-3, and velocity jumps in 20 % of the cells
    assert st.meta["n0"] == 1.0e17 and env.meta["n0"] == 1.0e14
    assert st.meta["jump_fraction"] == 0.2 and env.meta["jump_fraction"] == 0.0
