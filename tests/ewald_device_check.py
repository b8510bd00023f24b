"""Ewald option through the device-resident entry point, run in its own process by
test_gpu_ewald.py (torch is initialised before libtholedipole_hip.so is loaded, so that both use
the one HIP runtime torch brings).

tdp_execute_device examines the list build's counts at the solver's first host check and enqueues
the forces ahead of that check; tdp_execute checks the counts right after the build.  Both must give
the same results:
* the 36,864-site box, mixed and FP64, three device evaluations in a row against the host entry
  point on the same context;
* the same context then evaluating the box compressed by 15% (1.63x the density): its stored lists
  overflow, the late check finds it and the evaluation runs again with grown lists; the results
  and the stored entry counts must match a new context's.
"""
import os
import sys

import torch

torch.cuda.init()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from tholedipoleplugin_amd import _lib, systems  # noqa: E402

dev = torch.device("cuda:0")
stream = torch.cuda.Stream(dev)


def device_eval(h, pos, box):
    p = torch.tensor(np.ascontiguousarray(pos).reshape(-1), dtype=torch.float64, device=dev)
    f = torch.zeros(p.numel(), dtype=torch.float64, device=dev)
    e = torch.zeros(1, dtype=torch.float64, device=dev)
    with torch.cuda.stream(stream):
        h.execute_device(p.data_ptr(), box, f.data_ptr(), e.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    return e.item(), f.cpu().numpy().reshape(-1, 3)


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


ok = True
for precision, tol in (("double", 1e-12), ("mixed", 1e-9)):
    s = systems.config1()
    h = _lib.HipContextHandle(s.force, s.box, precision=precision, solver="pcg", flags=_lib.FLAG_EWALD)
    e_host, f_host = h.execute(s.positions, s.box)
    for k in range(3):
        e_dev, f_dev = device_eval(h, s.positions, s.box)
        de, df = abs(e_dev - e_host) / abs(e_host), rel(f_dev, f_host)
        print("%s device evaluation %d vs host entry point: energy %.3e, forces %.3e" % (precision, k, de, df),
              flush=True)
        ok &= de <= tol and df <= tol
    # compressed box on the same context: the lists sized for the first box overflow
    box2 = s.box * 0.85
    pos2 = s.positions * 0.85
    info1 = h.ewald_info()
    e2, f2 = device_eval(h, pos2, box2)
    info2 = h.ewald_info()
    h.close()
    h2 = _lib.HipContextHandle(s.force, box2, precision=precision, solver="pcg", flags=_lib.FLAG_EWALD)
    e_ref, f_ref = h2.execute(pos2, box2)
    info_ref = h2.ewald_info()
    h2.close()
    de, df = abs(e2 - e_ref) / abs(e_ref), rel(f2, f_ref)
    n1 = info1["entries32"] + info1["entries64"]
    n2 = info2["entries32"] + info2["entries64"]
    nr = info_ref["entries32"] + info_ref["entries64"]
    print("%s compressed box: entries %d -> %d (new context %d); energy %.3e, forces %.3e vs a new context"
          % (precision, n1, n2, nr, de, df), flush=True)
    ok &= n2 == nr and n2 > 1.5 * n1 and de <= tol and df <= tol
print("OK" if ok else "MISMATCH")
sys.exit(0 if ok else 1)
