"""find_unstable / apply_flip_soa (mgkmc/stz/cascade.py:4-113) on the GPU."""
from ..kmc_device import find_below_host, flip_host


def find_unstable(Q_field, threshold=0.0):
    """(count, 4) int64 (x, y, z, m) of modes with Q < threshold, in order."""
    return find_below_host(Q_field, threshold)


def apply_flip_soa(eps_plastic_field, key_field, soft_prop_field, last_event_time_field, catalog, x, y, z, m,
                   current_time, jp, jt, g_max, jn_frac):
    """Flip mode m of voxel (x, y, z), in place on the host arrays."""
    flip_host(3, eps_plastic_field, soft_prop_field, last_event_time_field, catalog, (x, y, z), m, current_time,
              jp, jt, g_max, jn_frac)
