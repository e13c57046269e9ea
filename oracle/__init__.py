"""CPU oracle for the equilibrium-solve hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``mgkmc_amd``) may
import, call or link anything in this directory.  The only users are
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``, where it is the checker / the timed CPU restatement, never the
thing that is measured or shipped.

Each function restates the reference algorithm (siamak-attarian/mgkmc) in
NumPy/SciPy and cites the reference file:line it follows.  The restatement
keeps the reference's memory structure (materialized Green operator,
per-component complex FFTs, sequential FP64 sums) so that it doubles as the
CPU baseline.

Pinning status (what checks each function against something other than its
own formula) is listed in the module docstrings and in tests/test_oracle_*.py.
"""
