"""2D KMC seams (mgkmc/kmc_simulator_functions.py:5-218) on the GPU."""
import numpy as np

from .kmc_device import barrier_host, rates_host, select_host, find_below_host, flip_host
from .stz import kmc as _k


def compute_rates_2d(Q_field, volume, temperature, nu0=1e13, cascade_mode=True):
    return rates_host(np.asarray(Q_field), volume, temperature, nu0, cascade_mode)


def select_event_2d(rates, total_rate):
    if total_rate <= 0:
        return -1
    r = total_rate * _k.SELECTION_STREAM.random_sample()
    return select_host(rates, r)


def decode_index_2d(flat_idx, ny, M):
    m = flat_idx % M
    t = flat_idx // M
    return t // ny, t % ny, m


def compute_barrier_2d(Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time, current_time,
                       prev_strain_dir, softening_cap, softening_scheme_idx=0, tau=np.inf):
    barrier_host(2, Q_field, Q0_field, sigma_field, catalog, volume, soft_prop, last_event_time, current_time,
                 prev_strain_dir, softening_cap, softening_scheme_idx, tau)


def find_unstable_2d(Q_field, threshold=0.0):
    return find_below_host(Q_field, threshold)


def apply_flip_soa_2d(eps_plastic_field, soft_prop_field, last_event_time_field, catalog, x, y, m, current_time,
                      jp, jt, g_max, jn_frac):
    flip_host(2, eps_plastic_field, soft_prop_field, last_event_time_field, catalog, (x, y), m, current_time, jp,
              jt, g_max, jn_frac)
