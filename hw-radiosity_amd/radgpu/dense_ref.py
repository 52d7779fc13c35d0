"""FP64 dense solve of B = E + diag(rho) F B on the GPU with torch (LU),
the reference the solvers' errors are measured against at sizes a CPU LU
cannot reach in a benchmark run.

Runs as its own process (torch ships its own HIP runtime, which does not
share a process with libradgpu's):

    python -m radgpu.dense_ref DIR

reads DIR/F.npy (n x n float32), DIR/rho.npy and DIR/E.npy (n x k float64)
and writes DIR/B.npy; columns with the same rho share one factorisation."""
import os
import sys
import time

import numpy as np


def main(d):
    import torch
    dev = torch.device("cuda:0")
    F = torch.from_numpy(np.load(os.path.join(d, "F.npy"), mmap_mode="r")[...]).to(dev).double()
    rho = np.load(os.path.join(d, "rho.npy"))
    E = np.load(os.path.join(d, "E.npy"))
    n = F.shape[0]
    B = np.zeros_like(E)
    groups = {}
    for c in range(rho.shape[1]):
        groups.setdefault(rho[:, c].tobytes(), []).append(c)
    t0 = time.perf_counter()
    eye = torch.eye(n, dtype=torch.float64, device=dev)
    for cols in groups.values():
        r = torch.from_numpy(rho[:, cols[0]]).to(dev)
        M = eye - r[:, None] * F
        X = torch.linalg.solve(M, torch.from_numpy(E[:, cols]).to(dev))
        B[:, cols] = X.cpu().numpy()
        del M, X
    torch.cuda.synchronize()
    np.save(os.path.join(d, "B.npy"), B)
    print("dense_ref: n=%d, %d factorisations, %.2f s" % (n, len(groups), time.perf_counter() - t0), flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
