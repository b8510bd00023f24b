"""Device-resident entry point check, run in its own process by test_gpu_parity.py.

torch is initialised before libtholedipole_hip.so is loaded, so that both use the one HIP
runtime torch brings (its libamdhip64.so.7 satisfies the library's dependency).
"""
import os
import sys

import torch

torch.cuda.init()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import ref_systems  # noqa: E402
from tholedipoleplugin_amd import _lib  # noqa: E402

rs = ref_systems.water4()
rs.force.setMutualInducedTargetEpsilon(1e-12)
rs.force.setMutualInducedMaxIterations(1000)
h = _lib.HipContextHandle(rs.force, rs.system.getDefaultPeriodicBoxVectors(), "double", "jacobi")
e_host, f_host = h.execute(rs.positions, rs.box)
dev = torch.device("cuda:0")
pos = torch.tensor(rs.positions.reshape(-1), dtype=torch.float64, device=dev)
frc = torch.ones(pos.numel(), dtype=torch.float64, device=dev)
en = torch.zeros(1, dtype=torch.float64, device=dev)
stream = torch.cuda.Stream(dev)
with torch.cuda.stream(stream):
    h.execute_device(pos.data_ptr(), rs.box, frc.data_ptr(), en.data_ptr(), stream.cuda_stream)
    total = frc.sum()   # ordered after the library's work on this stream
torch.cuda.synchronize()
f = frc.cpu().numpy().reshape(-1, 3) - 1.0
df = np.max(np.abs(f - f_host)) / np.max(np.abs(f_host))
de = abs(en.item() - e_host) / abs(e_host)
dt = abs(total.item() - (f_host.sum() + f.size)) / f.size
print("device entry: force max rel diff %.3e, energy rel diff %.3e, stream-ordered sum diff %.3e" % (df, de, dt))
ok = df <= 1e-12 and de <= 1e-12 and dt <= 1e-9
print("OK" if ok else "MISMATCH")
sys.exit(0 if ok else 1)
