"""Writes the oracle's results on two headline workloads as compact fixtures
for the GPU tests (tests/test_gpu_configs.py):

    python tests/golden/make_fixtures.py cfg0        # ~1-2 min
    python tests/golden/make_fixtures.py cfg1        # 500 events at 128^3: ~1-2 h on 8 cores

cfg0 -> cfg0_aqs_64.npz: the event list, (eps_macro, <sigma>) after each of
the 200 increments and the final stress and eigenstrain fields.
cfg1 -> cfg1_kmc_128.npz: the event list (x, y, z, m, time, is_instab), the
final macro strain and mean stress, norms of the final stress field, the
stress at 256 sampled voxels, the final eigenstrain sum and Q0 checksum.

Inputs come from cfg_inputs.py; the oracle is oracle/aqs_loop.py (fast mode at
128^3: matrix-free LS and planar barriers, pinned against the standard mode in
tests/test_oracle_kmc.py) and oracle/kmc2d_loop.py.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import cfg_inputs as ci  # noqa: E402


def cfg0(out):
    from oracle.kmc2d_loop import run_events_2d
    c = ci.CFG0
    E, nu, ep = ci.cfg0_inputs(c["n"])
    t0 = time.time()
    ev, st = run_events_2d(c["n"], c["n"], c["M"], c["gamma0"], E, nu, c["pixel"], c["temperature"], c["strain_rate"],
                           c["step_size"], (0, 0), 10 ** 9, seed=c["seed"], sel_seed=c["sel_seed"],
                           barrier_kwargs=ci.BARRIERS, stress_targets={(1, 1): 0.0}, mixed_tol=c["mixed_tol"],
                           ls_tol=c["ls_tol"], n_global_steps=c["steps"], eps_plastic=ep, **ci.STZ)
    print(f"cfg0: {len(ev)} events, {len(st['increments'])} increments in {time.time() - t0:.1f} s", flush=True)
    sig = st["sig"]
    np.savez_compressed(out, events=np.array([e[:3] for e in ev], dtype=np.int32).reshape(-1, 3),
                        times=np.array([e[3] for e in ev]), instab=np.array([e[4] for e in ev], dtype=bool),
                        inc_eps=np.array([e for e, _ in st["increments"]]),
                        inc_sig=np.array([s for _, s in st["increments"]]),
                        sig=np.stack([sig[..., 0, 0], sig[..., 1, 1], sig[..., 0, 1]]),
                        eps_plastic=np.stack([st["eps_plastic"][..., 0, 0], st["eps_plastic"][..., 1, 1],
                                              st["eps_plastic"][..., 0, 1]]),
                        eps_macro=st["eps_macro"])


def cfg1(out, events):
    from oracle import ls as ol
    from oracle.aqs_loop import run_events
    c = ci.CFG1
    n = c["n"]
    ol.FFT_WORKERS = len(os.sched_getaffinity(0))
    E, nu, ep = ci.cfg1_inputs(n)
    t0 = time.time()

    def progress(k):
        if k % 10 == 0:
            print(f"cfg1: {k} events, {time.time() - t0:.0f} s", flush=True)
    ev, st = run_events(n, n, n, c["M"], c["gamma0"], E, nu, c["pixel"], c["temperature"], c["strain_rate"],
                        c["step_size"], (0, 0), events, seed=c["seed"], sel_seed=c["sel_seed"],
                        barrier_kwargs=ci.BARRIERS, ls_tol=c["ls_tol"], mixed_tol=c["mixed_tol"], eps_plastic=ep,
                        fast=True, progress=progress, **ci.STZ)
    print(f"cfg1: {len(ev)} events in {time.time() - t0:.1f} s", flush=True)
    sig = st["sig"]
    pick = np.random.default_rng(123).choice(n ** 3, 256, replace=False)
    flat = sig.reshape(-1, 3, 3)
    np.savez_compressed(out, events=np.array([e[:4] for e in ev], dtype=np.int32).reshape(-1, 4),
                        times=np.array([e[4] for e in ev]), instab=np.array([e[5] for e in ev], dtype=bool),
                        eps_macro=st["eps_macro"], sig_mean=sig.mean(axis=(0, 1, 2)),
                        sig_norm=np.sqrt(np.sum(sig ** 2)), sig_absmax=np.abs(sig).max(),
                        sample_index=pick, sample_sig=flat[pick], eps_plastic_sum=st["eps_plastic"].sum(axis=(0, 1, 2)),
                        eps_plastic_norm=np.sqrt(np.sum(st["eps_plastic"] ** 2)), q0_sum=st["Q0"].sum(),
                        time=st["time"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["cfg0", "cfg1"])
    ap.add_argument("--events", type=int, default=ci.CFG1["events"])
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.which == "cfg0":
        cfg0(a.out or os.path.join(HERE, "cfg0_aqs_64.npz"))
    else:
        cfg1(a.out or os.path.join(HERE, "cfg1_kmc_128.npz"), a.events)


if __name__ == "__main__":
    main()
