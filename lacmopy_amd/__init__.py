"""lacmopy_amd: gfx950-native super-droplet (SIP) time step of LacmoPy.

Python mirror of the reference entry points (integration.simulate and step
functions, collision.all_or_nothing routines) over hand-written HIP kernels
(liblacmo_amd.so, C ABI in include/lacmo_amd.h).
"""
import os

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_kernel_data_Ecol(kernel_type="Long_Bott"):
    """E_col table and log-radius grid parameters (file_handling.py:764-804)."""
    import numpy as np
    d = os.path.join(DATA, "Ecol_grid_data", kernel_type)
    radius_grid = np.load(os.path.join(d, "radius_grid_out.npy"))
    E_col_grid = np.load(os.path.join(d, "E_col_grid.npy"))
    R_kernel_low_log = np.log(radius_grid[0])
    bin_factor_R_log = np.log(radius_grid[1] / radius_grid[0])
    return E_col_grid, R_kernel_low_log, bin_factor_R_log, len(radius_grid)
