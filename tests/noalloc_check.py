"""Allocation-free entry point check, run in its own process by the no-alloc GPU test.

tdp_execute_device_noalloc against tdp_execute_device through the C ABI, on two calls in a row,
the second with a smaller system: a 3,000-site water cluster (NoCutoff, mixed precision, PCG),
then the 648-site water box (reference PME semantics, FP64, PCG).  Each is compared with a twin
context evaluated by tdp_execute_device (which computes its own spatial order: in FP64 the sums
differ in order only, checked at 1e-10; in mixed precision the order also decides which tile pairs
run in FP32) at the parity bar; then the noalloc entry point on the twin context
itself (same order) must give bit-identical results; a context too large for the scratch is
refused with an error before any work.  torch owns the device buffers and the stream.
"""
import os
import sys

import torch

torch.cuda.init()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

from tholedipoleplugin_amd import _lib, systems  # noqa: E402

dev = torch.device("cuda:0")
stream = torch.cuda.Stream(dev)
E_TOL, F_TOL = 1e-6, 1e-5


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def run(h, pos, box, noalloc):
    d_pos = torch.tensor(pos.reshape(-1), dtype=torch.float64, device=dev)
    d_f = torch.zeros(pos.size, dtype=torch.float64, device=dev)
    d_e = torch.zeros(1, dtype=torch.float64, device=dev)
    with torch.cuda.stream(stream):
        if noalloc:
            h.execute_device_noalloc(d_pos.data_ptr(), box, d_f.data_ptr(), d_e.data_ptr(), stream.cuda_stream)
        else:
            h.execute_device(d_pos.data_ptr(), box, d_f.data_ptr(), d_e.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    return d_e.item(), d_f.cpu().numpy().reshape(-1, 3)


cl = systems.water_cluster(1000, seed=4242)
cl.force.setMutualInducedTargetEpsilon(1e-9)
bx = systems.config0()
bx.force.setMutualInducedTargetEpsilon(1e-9)
cases = [("cluster3000 mixed", cl.force, np.asarray(cl.positions, np.float64), None, np.zeros(9), "mixed"),
         ("box648 double", bx.force, np.asarray(bx.positions, np.float64), bx.box, bx.box, "double")]
ok = True
refs = []
for name, f, pos, box, dbox, prec in cases:
    twin = _lib.HipContextHandle(f, dbox, precision=prec, solver="pcg")
    refs.append((twin, run(twin, pos, box, noalloc=False)))
handles = []
for k, (name, f, pos, box, dbox, prec) in enumerate(cases):
    print("noalloc call %d: %s, %d sites" % (k + 1, name, pos.shape[0]), flush=True)
    h = _lib.HipContextHandle(f, dbox, precision=prec, solver="pcg")
    handles.append(h)
    e, fr = run(h, pos, box, noalloc=True)
    e_ref, f_ref = refs[k][1]
    de, df = abs(e - e_ref) / abs(e_ref), rel_rms(fr, f_ref)
    st = h.solver_stats()
    print("  energy rel %.3e, force rel RMS %.3e, iterations %d, converged %d" % (de, df, st["iterations"],
                                                                                 st["converged"]), flush=True)
    # FP64: only the summation order differs; mixed: the internal order also decides which tile pairs
    # run in FP32, so the results differ at the FP32 level, within the bar
    tight = prec == "mixed" or (de <= 1e-10 and df <= 1e-10)
    ok &= de <= E_TOL and df <= F_TOL and tight and st["converged"] == 1
for k, (name, f, pos, box, dbox, prec) in enumerate(cases):
    twin, (e_ref, f_ref) = refs[k]
    e, fr = run(twin, pos, box, noalloc=True)
    same = e == e_ref and np.array_equal(fr, f_ref)
    print("same order, %s: bit-identical %s" % (name, same), flush=True)
    ok &= same
big = systems.water_cluster(3000, seed=4343)
hb = _lib.HipContextHandle(big.force, np.zeros(9), precision="double", solver="pcg")
d_pos = torch.tensor(np.asarray(big.positions).reshape(-1), dtype=torch.float64, device=dev)
rc = hb.execute_device_noalloc(d_pos.data_ptr(), None, 0, 0, stream.cuda_stream, check=False)
msg = hb.L.tdp_last_error(hb.h).decode()
print("9,000 sites: status %d (%s)" % (rc, msg), flush=True)
ok &= rc == _lib.TDP_ERR_INVALID_ARGUMENT and "scratch" in msg
print("OK" if ok else "MISMATCH")
sys.exit(0 if ok else 1)
