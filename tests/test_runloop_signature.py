"""The simulators' run_simulation take the reference's parameters in the
reference's order (mgkmc/aqs.py:682-705, mgkmc/kmc_simulator.py:796-803), with
the one extension max_events appended last; the log writers keep the
reference's fixed-width formats (mgkmc/aqs.py:377-415)."""
import inspect

import numpy as np

# mgkmc/aqs.py:682-705
REF_3D = ["self", "n_global_steps", "step_size", "component", "stress_targets", "mixed_tol", "mixed_max_iter",
          "checkpoint_interval", "checkpoint_path", "stop_on_stress_drop", "stress_drop_component",
          "stop_post_drop_steps", "vtk_interval", "vtk_elastic_only", "track_cascades", "enable_console_log",
          "summary_filename", "enable_summary_log", "enable_global_log", "enable_cascade_log", "enable_kmc_log",
          "checkpoint_elastic_only", "enable_save_q", "save_q_interval", "save_q_elastic_only", "max_kmc_steps_pct",
          "max_cascade_steps_pct", "ignore_drop_steps", "stress_drop_lookback", "append_logs", "eps_target",
          "cascade_mode"]
REF_3D_DEFAULTS = {"component": (0, 1), "mixed_tol": 1e-4, "mixed_max_iter": 50, "checkpoint_interval": None,
                   "checkpoint_path": "checkpoint", "stop_on_stress_drop": None, "stop_post_drop_steps": 20,
                   "vtk_interval": "none", "enable_cascade_log": True, "enable_kmc_log": True,
                   "max_kmc_steps_pct": 0.3, "ignore_drop_steps": 0, "stress_drop_lookback": 1, "eps_target": None}
# mgkmc/kmc_simulator.py:796-803
REF_2D = ["self", "n_global_steps", "step_size", "component", "stress_targets", "mixed_tol", "mixed_max_iter",
          "checkpoint_interval", "checkpoint_path", "vtk_interval", "vtk_elastic_only", "track_cascades",
          "enable_console_log", "summary_filename", "enable_summary_log", "enable_global_log", "max_kmc_steps_pct",
          "max_cascade_steps_pct", "enable_kmc_log"]
REF_2D_DEFAULTS = {"component": (0, 0), "checkpoint_interval": "none", "enable_kmc_log": False}


def _params(fn):
    return list(inspect.signature(fn).parameters.values())


def test_run_simulation_3d_signature():
    from mgkmc_amd.aqs import ThermalSimulation
    ps = _params(ThermalSimulation.run_simulation)
    names = [p.name for p in ps]
    assert names[: len(REF_3D)] == REF_3D
    assert names[len(REF_3D):] == ["max_events", "kwargs"]
    for p in ps:
        if p.name in REF_3D_DEFAULTS:
            assert p.default == REF_3D_DEFAULTS[p.name], p.name


def test_run_simulation_2d_signature():
    from mgkmc_amd.kmc_simulator import KmcSimulation2D
    ps = _params(KmcSimulation2D.run_simulation)
    names = [p.name for p in ps]
    assert names[: len(REF_2D)] == REF_2D
    assert names[len(REF_2D):] == ["max_events", "kwargs"]
    for p in ps:
        if p.name in REF_2D_DEFAULTS:
            assert p.default == REF_2D_DEFAULTS[p.name], p.name


def test_log_formats(tmp_path):
    from mgkmc_amd.runlog import RunLog, StressDrop
    log = RunLog(str(tmp_path), 3, enable_console_log=False)
    eps = np.zeros((3, 3))
    eps[0, 0] = 0.0123456789
    sig = np.diag([1.23456e9, 0.0, -2.5e8])
    log.log_global(7, 2, 5, 1.5e-9, eps, sig, 0, 1)
    log.summary(7, "kmc", eps[0, 0], sig[0, 0], 5, 0, 1, 1.5e-9)
    log.close()
    g = (tmp_path / "global_log.txt").read_text().splitlines()
    assert g[0].split() == ["GlobalStep", "ElasticStep", "KMCStep", "Eps_xx", "Eps_yy", "Eps_zz", "Eps_xy",
                            "Eps_xz", "Eps_yz", "Sig_xx(GPa)", "Sig_yy(GPa)", "Sig_zz(GPa)", "Sig_xy(GPa)",
                            "Sig_xz(GPa)", "Sig_yz(GPa)", "CascadeSteps", "TotalCascadeFlips", "SimTime(s)"]
    # "{:<10d} {:<12d} {:<10d} " then 6 x "{:<15.6f}", 6 x "{:<15.3f}", "{:<14d} {:<17d} {:<15.6e}"
    assert g[1].startswith("7          2            5          0.012346        ")
    f = g[1].split()
    assert f[9] == "1.235" and f[11] == "-0.250" and f[-1] == "1.500000e-09"
    sm = (tmp_path / "summary_log.txt").read_text().splitlines()
    assert sm[0].split()[:4] == ["Timestamp", "Elapsed(s)", "Step", "Type"]
    assert set(sm[1]) == {"-"} and len(sm[1]) == len(sm[0]) + 1
    assert sm[2].split()[3:] == ["7", "KMC", "0.012346", "1.235", "5", "0", "1", "1.500000e-09"]
    assert (tmp_path / "cascade_log.txt").exists() is False  # enable_cascade_log defaults off in RunLog
    # stress drop: lookback 1 compares the increment with itself (never fires); lookback 2 with the entry before
    d1 = StressDrop(0.1, 0, 0, 1, 1.0)
    d1.after_increment(1, 0.5)
    assert not d1.triggered
    d2 = StressDrop(0.1, 1, 0, 2, 1.0)
    d2.after_increment(1, 0.5)
    assert d2.triggered and d2.trigger_step == 1
    assert d2.stop_now() is False and d2.stop_now() is True
