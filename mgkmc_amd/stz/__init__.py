"""STZ / KMC seams (mgkmc/stz/) on the GPU."""
