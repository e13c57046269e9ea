"""update_stress_fft_full (mgkmc/stz/update_fft.py:4-22) on the GPU."""
from ..linear_elastic_simulator import spectral_solver_3d


def update_stress_fft_full(eps_plastic_field, eps_macro, E, nu, pixel=1.0,
                           max_iter=200, tol=1e-6, verbose=False):
    """Cold-start 3D LS solve with eps_plastic; returns (eps, sig, <eps>, <sig>)."""
    return spectral_solver_3d(E, nu, eps_macro, eps_plastic=eps_plastic_field,
                              max_iter=max_iter, tol=tol, verbose=verbose, pixel=pixel)
