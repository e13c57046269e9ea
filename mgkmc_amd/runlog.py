"""Fixed-width text logs and the stress-drop stop of the simulators' run loops.

The device simulators write the reference's logs with the reference's
columns and formats:

  * 3D (mgkmc/aqs.py:353-415, :750-778): ``summary_log.txt`` (or the given
    summary file name), ``global_log.txt``, ``cascade_log.txt`` (header only:
    the reference never writes a cascade line) and ``kmc_log.txt`` (opened
    lazily by log_kmc, which the reference's loop never calls, so the file is
    not created);
  * 2D (mgkmc/kmc_simulator.py:177-224, :535-557, :821-843):
    ``summary_log.txt``, ``global_log.txt`` with the plane-strain sigma_zz
    column, and ``flipped_voxels_log.txt`` (one line per flipped mode, with
    its softening and barrier at selection) when enable_kmc_log is set.

``StressDrop`` is the reference's early stop (mgkmc/aqs.py:1151-1181).
Timestamps and elapsed wall times are this run's.
"""
import os
import time
from datetime import datetime

import numpy as np

SUMMARY_HEADER = (f"{'Timestamp':<22} {'Elapsed(s)':<12} {'Step':<8} {'Type':<15} {'Eps_xx':<12} "
                  f"{'Sig_xx(GPa)':<15} {'KMC':<8} {'Cascade':<8} {'Flips':<8} {'SimTime(s)':<15}")
GLOBAL_HEADERS = ["GlobalStep", "ElasticStep", "KMCStep", "Eps_xx", "Eps_yy", "Eps_zz", "Eps_xy", "Eps_xz", "Eps_yz",
                  "Sig_xx(GPa)", "Sig_yy(GPa)", "Sig_zz(GPa)", "Sig_xy(GPa)", "Sig_xz(GPa)", "Sig_yz(GPa)",
                  "CascadeSteps", "TotalCascadeFlips", "SimTime(s)"]
GLOBAL_HEADER_FMT = "{:<10} {:<12} {:<10} " + " ".join(["{:<15}"] * 12) + " {:<14} {:<17} {:<15}"
GLOBAL_FMT = ("{:<10d} {:<12d} {:<10d} " + " ".join(["{:<15.6f}"] * 6) + " " + " ".join(["{:<15.3f}"] * 6) +
              " {:<14d} {:<17d} {:<15.6e}\n")
FLIP_HEADER = f"{'Step':<10} {'X':<6} {'Y':<6} {'Mode':<6} {'softening':<14} {'energy barrier':<18} {'Type':<10}\n"


def _empty(path):
    return not os.path.exists(path) or os.path.getsize(path) == 0


class RunLog:
    """The run loop's log files and console lines for one run_simulation call.

    ``ndim`` selects the 3D or 2D formats; ``thermal`` adds the T_avg column
    (its value from ``t_avg()``)."""

    def __init__(self, output_dir, ndim, summary_filename="summary_log.txt", enable_summary_log=True,
                 enable_global_log=True, enable_cascade_log=False, enable_console_log=True, append=False,
                 thermal=False, t_avg=None, plane_mode="plane_strain", nu_avg=0.0):
        self.ndim = ndim
        self.thermal = thermal
        self.t_avg = t_avg
        self.console = enable_console_log
        self.plane_mode = plane_mode
        self.nu_avg = nu_avg
        self.t0 = time.time()
        mode = "a" if append else "w"
        self.summary_path = os.path.join(output_dir, summary_filename)
        self.global_path = os.path.join(output_dir, "global_log.txt")
        self.cascade_path = os.path.join(output_dir, "cascade_log.txt")
        self.f_summary = self.f_global = self.f_cascade = None
        header = SUMMARY_HEADER + (" T_avg(K)" if thermal else "") + "\n"
        self.header = header
        if ndim == 3:
            # mgkmc/aqs.py:369-397: files opened first, headers when new or truncated
            self.f_global = open(self.global_path, mode, buffering=1) if enable_global_log else None
            self.f_cascade = open(self.cascade_path, mode, buffering=1) if enable_cascade_log else None
            self.f_summary = open(self.summary_path, mode, buffering=1) if enable_summary_log else None
            if self.f_summary and (not append or _empty(self.summary_path)):
                self.f_summary.write(header)
                self.f_summary.write("-" * len(header) + "\n")
            if self.f_global and (not append or _empty(self.global_path)):
                self.f_global.write(GLOBAL_HEADER_FMT.format(*GLOBAL_HEADERS) + "\n")
            if self.f_cascade and (not append or _empty(self.cascade_path)):
                self.f_cascade.write("{:<12} {:<12} {:<15} {:<30}".format(
                    "GlobalStep", "LocalStep", "NumUnstable", "FlippedVoxels(x,y,z,mode)") + "\n")
        else:
            # mgkmc/kmc_simulator.py:177-205
            if enable_summary_log:
                self.f_summary = open(self.summary_path, mode, buffering=1)
                if not append or _empty(self.summary_path):
                    self.f_summary.write(header)
                    self.f_summary.write("-" * len(header) + "\n")
            if enable_global_log:
                self.f_global = open(self.global_path, mode, buffering=1)
                if not append or _empty(self.global_path):
                    self.f_global.write(GLOBAL_HEADER_FMT.format(*GLOBAL_HEADERS) + "\n")

    def console_header(self):
        """The console header the 3D loop prints before the run (aqs.py:731-736)."""
        if self.console:
            h = SUMMARY_HEADER + (" T_avg(K)" if self.thermal else "")
            print(h)
            print("-" * len(h))

    def log_global(self, step, elastic_steps, kmc_steps, sim_time, eps, sig, cascade_steps, total_flips):
        if not self.f_global:
            return
        if self.ndim == 3:
            data = [step, elastic_steps, kmc_steps, eps[0, 0], eps[1, 1], eps[2, 2], eps[0, 1], eps[0, 2], eps[1, 2],
                    sig[0, 0] / 1e9, sig[1, 1] / 1e9, sig[2, 2] / 1e9, sig[0, 1] / 1e9, sig[0, 2] / 1e9,
                    sig[1, 2] / 1e9, cascade_steps, total_flips, sim_time]
        else:
            # plane strain: sigma_zz = nu_avg (sigma_xx + sigma_yy) (kmc_simulator.py:211-224)
            szz = self.nu_avg * (sig[0, 0] + sig[1, 1]) if self.plane_mode == "plane_strain" else 0.0
            data = [step, elastic_steps, kmc_steps, eps[0, 0], eps[1, 1], 0.0, eps[0, 1], 0.0, 0.0,
                    sig[0, 0] / 1e9, sig[1, 1] / 1e9, szz / 1e9, sig[0, 1] / 1e9, 0.0, 0.0,
                    cascade_steps, total_flips, sim_time]
        self.f_global.write(GLOBAL_FMT.format(*data))

    def summary(self, step, step_type, strain, stress, kmc_steps, cascade_id, flips, sim_time, first_2d=False):
        """One summary line (aqs.py:762-775, kmc_simulator.py:825-843)."""
        now = datetime.now().strftime("%Y-%m-%d %H:%M:%S")
        elapsed = time.time() - self.t0
        line = (f"{now:<22} {elapsed:<12.2f} {step:<8d} {step_type.upper():<15} {strain:<12.6f} "
                f"{stress / 1e9:<15.3f} {kmc_steps:<8d} {cascade_id:<8d} {flips:<8d} {sim_time:<15.6e}")
        if self.thermal:
            line += f" {self.t_avg():<15.2f}"
        line += "\n"
        if self.f_summary:
            self.f_summary.write(line)
        if self.console:
            if first_2d:
                h = SUMMARY_HEADER + (" T_avg(K)" if self.thermal else "")
                print(h)
                print("-" * len(h))
            print(line.strip())

    def finish(self):
        """Duration line of a completed run (aqs.py:1193-1200, kmc_simulator.py:1148-1155)."""
        total = time.time() - self.t0
        m, s = divmod(total, 60)
        h, m = divmod(m, 60)
        msg = (f"\nSimulation Finish Time: {datetime.now().strftime('%Y-%m-%d %H:%M:%S')}\n"
               f"Total Duration: {total:.2f} seconds ({int(h):d}h {int(m):02d}m {int(s):02d}s)\n")
        if self.f_summary:
            self.f_summary.write(msg)
        print(msg)

    def close(self):
        for name in ("f_global", "f_cascade", "f_summary"):
            f = getattr(self, name)
            if f:
                f.close()
                setattr(self, name, None)


class FlipLog:
    """flipped_voxels_log.txt of the 2D simulator (kmc_simulator.py:535-557,
    :817-819): header at the start of a run, then one line per flip."""

    def __init__(self, output_dir):
        self.path = os.path.join(output_dir, "flipped_voxels_log.txt")
        with open(self.path, "w") as f:
            f.write(FLIP_HEADER)

    def write(self, step, x, y, m, softening, barrier, flip_type):
        with open(self.path, "a") as f:
            f.write(f"{step:<10d} {x:<6d} {y:<6d} {m:<6d} {softening:<14.6f} {barrier:<18.6f} {flip_type:<10}\n")


class StressDrop:
    """Early stop on a stress drop (mgkmc/aqs.py:1151-1181): the history holds
    the driven stress after the initial relaxation, after every KMC event and
    after every elastic increment; after an increment with step >
    ignore_drop_steps the drop (|s_ref| - |s|) / |s_ref| against the entry
    ``lookback`` back (the current one for lookback 1) above the threshold
    arms the stop, which ends the run after stop_post_drop_steps further
    increments."""

    def __init__(self, threshold, post_steps, ignore_steps, lookback, first):
        self.threshold = threshold
        self.countdown = post_steps
        self.ignore_steps = ignore_steps
        self.lookback = max(1, lookback)
        self.history = [first]
        self.triggered = False
        self.trigger_step = None

    def add(self, value):
        self.history.append(value)

    def after_increment(self, step, value):
        """Record the increment's stress; returns True when the run stops here."""
        self.history.append(value)
        if self.threshold is not None and not self.triggered and step > self.ignore_steps:
            if len(self.history) >= self.lookback:
                ref = self.history[-self.lookback]
                drop = (abs(ref) - abs(value)) / abs(ref) if abs(ref) > 1e-6 else 0.0
                if drop > self.threshold:
                    print(f"\n[ALERT] Stress Drop Detected! {drop * 100:.1f}% at step {step}")
                    self.triggered = True
                    self.trigger_step = step
        return False

    def stop_now(self):
        """The countdown after the increment's checkpoint / Q output (aqs.py:1179-1181)."""
        if self.triggered:
            if self.countdown > 0:
                self.countdown -= 1
                return False
            return True
        return False


def save_q(output_dir, step, Q):
    """Q of the current state, Q_step_XXXXXX.npy (aqs.py:1174-1177)."""
    np.save(os.path.join(output_dir, f"Q_step_{step:06d}.npy"), Q)
