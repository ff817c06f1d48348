"""Writes tests/golden/random_blocks_30q_seed42.json with the CPU oracle.

The at-scale plain workload, 30 qubits complex128 (16 GiB state), depth 40,
circuit seed 42 and measurement seed 42 (``workloads.random_blocks``): 2440
gates, 8 measurements, 8 resets and 4 conditional x.  The oracle's OpenMP
restatement of the reference's Rayon kernels (src/engine.rs:24-257, the
dispatcher of src/monolithic/statevector/simulator.rs:386-632) runs it from
|0..0> with ChaCha8Rng::seed_from_u64(42); the fixture keeps

* ``classical_bits``: the measurement record;
* ``p1``: the p1 every measurement and reset saw, in program order;
* ``amp_index`` / ``amp_re`` / ``amp_im``: 1024 amplitudes of the final state
  at indices drawn by numpy default_rng(2024) (sorted), plus 0 and 2^30-1;
* ``marg_low8`` / ``marg_high8``: probabilities() over qubits [7..0] and
  [29..22] (256 bins each; every amplitude contributes to one bin of each).

Needs about 17 GiB of host RAM and runs for tens of minutes on 8 cores:
``python tests/golden/make_random_blocks_30q.py``.  Every float is written
with ``repr``, which round-trips exactly.
"""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from dqsim_bosonic_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

N, DEPTH, SEED = 30, 40, 42
OUT = os.path.join(HERE, f"random_blocks_{N}q_seed{SEED}.json")


def sample_indices(n: int) -> np.ndarray:
    rng = np.random.default_rng(2024)
    idx = rng.choice(1 << n, size=1024, replace=False)
    return np.unique(np.concatenate([idx, [0, (1 << n) - 1]])).astype(np.int64)


def main(n: int = N, depth: int = DEPTH, out: str = OUT, ckpt: str | None = None) -> None:
    """``ckpt``: a directory for a resumable run; the state, record and RNG
    buffer are saved there every 300 instructions."""
    circ = W.random_blocks(n, depth, seed=SEED)
    text = circ.model_dump_json()
    insts = json.loads(text)["instructions"]
    sim = O.OracleSimulator(seed=SEED, parallel=True)
    cbits: dict[int, int] = {}
    rng = O.Rng(SEED)
    p1s: list[float] = []
    start = 0
    meta_path = os.path.join(ckpt, "meta.json") if ckpt else None
    if meta_path and os.path.exists(meta_path):
        with open(meta_path) as f:
            meta = json.load(f)
        state = np.load(os.path.join(ckpt, "state.npy"), mmap_mode=None)
        cbits = {int(k): v for k, v in meta["cbits"].items()}
        p1s = meta["p1"]
        start = meta["next"]
        rng._buf.raw = bytes.fromhex(meta["rng"])
        print(f"resumed at instruction {start}", flush=True)
    else:
        state = np.zeros(1 << n, dtype=np.complex128)
        state[0] = 1.0
    t0 = time.time()
    for i in range(start, len(insts)):
        inst = insts[i]
        sim._run(state, n, inst, cbits, rng, p1s)
        if i % 61 == 0 or inst["kind"] in ("measure", "reset"):
            print(f"[{time.time() - t0:8.1f} s] {i + 1}/{len(insts)} {inst['kind']} p1s={len(p1s)}", flush=True)
        if ckpt and (i + 1) % 300 == 0 and i + 1 < len(insts):
            os.makedirs(ckpt, exist_ok=True)
            np.save(os.path.join(ckpt, "state.tmp.npy"), state)
            os.replace(os.path.join(ckpt, "state.tmp.npy"), os.path.join(ckpt, "state.npy"))
            with open(meta_path + ".tmp", "w") as f:
                json.dump({"next": i + 1, "cbits": {str(k): v for k, v in cbits.items()}, "p1": p1s,
                           "rng": rng._buf.raw.hex()}, f)
            os.replace(meta_path + ".tmp", meta_path)
            print(f"[{time.time() - t0:8.1f} s] checkpoint at {i + 1}", flush=True)
    idx = sample_indices(n)
    amps = state[idx]
    marg_low = O.marginal_probs(state, n, list(range(7, -1, -1)))
    marg_high = O.marginal_probs(state, n, list(range(n - 1, n - 9, -1)))
    fix = {
        "circuit": {"generator": "workloads.random_blocks", "n": n, "depth": depth, "seed": SEED},
        "measure_seed": SEED,
        "classical_bits": {str(k): int(v) for k, v in sorted(cbits.items())},
        "p1": [float(p) for p in p1s],
        "amp_index": [int(i) for i in idx],
        "amp_re": [float(a.real) for a in amps],
        "amp_im": [float(a.imag) for a in amps],
        "marg_low8": [float(p) for p in marg_low],
        "marg_high8": [float(p) for p in marg_high],
        "oracle_seconds": round(time.time() - t0, 1),
    }
    with open(out, "w") as f:
        json.dump(fix, f, indent=0)
    print(f"wrote {out} in {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    # python make_random_blocks_30q.py [n out.json] [--ckpt DIR]
    args = sys.argv[1:]
    ck = None
    if "--ckpt" in args:
        ck = args[args.index("--ckpt") + 1]
        del args[args.index("--ckpt"):args.index("--ckpt") + 2]
    if args:
        main(int(args[0]), DEPTH, args[1], ck)
    else:
        main(ckpt=ck)
