"""Command-line access to the reference solver built from source
(oracle/_ref/libpearl_ref.so), for processes that also use the GPU: the
reference library carries its own HDF5/zlib and C++ runtime, so it is only
ever loaded in a separate process started from here.

TEST INFRASTRUCTURE ONLY (tests/ and the CPU-baseline leg of bench.py).

  python oracle/ref_tool.py read-checkpoint INPUT.in FRAME.h5 OUT.npz
      load a checkpoint with the reference reader (signature, closure and
      ledger checks included) and save phi, carbon, mu, scalars
  python oracle/ref_tool.py sharp-labels INPUT.in OUT.npz
      the phases of the reference's initialize at its init_sharp observer
      call (the FILL labels before smoothing)
  python oracle/ref_tool.py initialize INPUT.in OUT.npz
      the reference's whole initialize (labels, smoothing, carbon set-up):
      phi, carbon, mu, scalars
  python oracle/ref_tool.py baseline INPUT.in STATE.npz SECONDS OUT.json
      time the reference advance() from the saved state, serially
  python oracle/ref_tool.py baseline-mpi INPUT.in STATE.npz NPROCS STEPS OUT.json
      write the saved state as a reference checkpoint (the reference's own
      writer), run the unmodified pearl_parallel under mpiexec on NPROCS
      ranks from it for STEPS accepted steps, and report steps and the step
      loop's wall time from its diagnostics.csv (wall_s of the snapshot
      after the last step minus that of the snapshot before the first)
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def call(*args, timeout=600):
    """Run a subcommand in a child process; raises on failure."""
    r = subprocess.run([sys.executable, os.path.join(HERE, "ref_tool.py"), *map(str, args)],
                       capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"ref_tool {args[0]} failed: {r.stderr.strip()[-2000:]}")
    return r.stdout


def mpiexec():
    for cand in ("/opt/conda/bin/mpiexec", "mpiexec"):
        if os.path.isabs(cand) and os.path.exists(cand):
            return cand
    import shutil
    return shutil.which("mpiexec")


def parallel_program():
    p = os.path.join(HERE, "_ref", "pearl_parallel")
    return p if os.path.exists(p) else None


def grid_split(nprocs, nx, ny):
    """px*py = nprocs with px | nx and py | ny, as square as possible."""
    best = None
    for px in range(1, nprocs + 1):
        if nprocs % px:
            continue
        py = nprocs // px
        if nx % px or ny % py:
            continue
        score = abs((nx / px) - (ny / py))
        if best is None or score < best[0]:
            best = (score, px, py)
    if best is None:
        raise ValueError(f"no {nprocs}-rank split of {nx}x{ny}")
    return best[1], best[2]


def read_checkpoint(inp, h5, tmpdir):
    out = os.path.join(tmpdir, os.path.basename(h5) + ".npz")
    call("read-checkpoint", inp, h5, out)
    z = np.load(out)
    return z["phi"], z["carbon"], z["mu"], z["scalars"]


def sharp_labels(inp, tmpdir):
    """phi[3, ny, nx] of the reference's FILL labels for input file `inp`."""
    out = os.path.join(tmpdir, os.path.basename(inp) + ".labels.npz")
    call("sharp-labels", inp, out)
    return np.load(out)["phi"]


def initialize(inp, tmpdir):
    """phi, carbon, mu, scalars of the reference's initialize for `inp`."""
    out = os.path.join(tmpdir, os.path.basename(inp) + ".init.npz")
    call("initialize", inp, out)
    z = np.load(out)
    return z["phi"], z["carbon"], z["mu"], z["scalars"]


def main(argv):
    sys.path.insert(0, HERE)
    import oracle as O
    cmd = argv[0]
    if cmd == "read-checkpoint":
        inp, h5, out = argv[1:4]
        R = O.Ref(inp)
        R.read_checkpoint(h5)
        phi, c, mu, sc = R.get_state()
        np.savez(out, phi=phi, carbon=c, mu=mu, scalars=sc)
        return 0
    if cmd == "sharp-labels":
        inp, out = argv[1:3]
        np.savez(out, phi=O.Ref(inp).sharp_labels())
        return 0
    if cmd == "initialize":
        inp, out = argv[1:3]
        R = O.Ref(inp)
        R.initialize()
        phi, c, mu, sc = R.get_state()
        np.savez(out, phi=phi, carbon=c, mu=mu, scalars=sc)
        return 0
    if cmd == "baseline":
        inp, state, seconds, out = argv[1], argv[2], float(argv[3]), argv[4]
        z = np.load(state)
        R = O.Ref(inp)
        R.set_state(z["phi"], z["carbon"], z["mu"], z["scalars"])
        stop = float(z["stop_time"])
        steps = 0
        t0 = time.perf_counter()
        while True:
            info, err = R.advance(stop)
            if err:
                raise RuntimeError(err)
            steps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                break
        with open(out, "w") as f:
            json.dump({"steps": steps, "seconds": el}, f)
        return 0
    if cmd == "baseline-mpi":
        import tempfile
        inp, state, nprocs, steps, out = argv[1], argv[2], int(argv[3]), int(argv[4]), argv[5]
        z = np.load(state)
        R = O.Ref(inp)
        R.set_state(z["phi"], z["carbon"], z["mu"], z["scalars"])
        tmp = argv[6] if len(argv) > 6 else tempfile.mkdtemp()  # keeps start.h5 and run/ there
        os.makedirs(tmp, exist_ok=True)
        ckpt = os.path.join(tmp, "start.h5")
        R.write_frame(ckpt)
        ny, nx = z["carbon"].shape
        px, py = grid_split(nprocs, nx, ny)
        run_dir = os.path.join(tmp, "run")
        cmd = [mpiexec(), "-n", str(nprocs), parallel_program(), inp, run_dir, "--restart", ckpt,
               "--max-steps", str(steps), "--px", str(px), "--py", str(py)]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(f"pearl_parallel failed: {r.stdout[-1000:]} {r.stderr[-2000:]}")
        with open(os.path.join(run_dir, "diagnostics.csv")) as f:
            rows = [line.strip().split(",") for line in f if line.strip()]
        head, first, last = rows[0], rows[1], rows[-1]
        si, wi = head.index("step"), head.index("wall_s")
        with open(out, "w") as f:
            json.dump({"steps": int(last[si]) - int(first[si]), "seconds": float(last[wi]) - float(first[wi]),
                       "nprocs": nprocs, "px": px, "py": py}, f)
        return 0
    raise SystemExit(f"unknown command {cmd}")


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
