"""compute_barrier (mgkmc/stz/barriers.py:4-100) on the GPU."""
import numpy as np

from ..kmc_device import barrier_host


def compute_barrier(Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time, current_time,
                    prev_strain_dir, softening_cap, softening_scheme_idx=0, tau=np.inf):
    """Barriers Q = Q0 exp(-g_eff) - W for all voxels and modes, in place on
    Q_field (nx, ny, nz, M); same arguments as the reference."""
    barrier_host(3, Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time, current_time,
                 prev_strain_dir, softening_cap, softening_scheme_idx, tau)
