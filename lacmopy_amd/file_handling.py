"""State files in the reference's formats (file_handling.py of LacmoPy).

Host-side I/O on either side of the time step: the file names, array
orders and text layouts are the reference's, so states can be handed back and
forth (spin-up hand-off, frame dumps read by the reference's evaluation code).
All files are plain .npy / text, written and read with numpy (no pickles).
"""
import os
from datetime import datetime

import numpy as np

from .grid import Grid


def _t(t):
    return str(int(t))


# ------------------------------------------------------------ parameters
def save_sim_paras_to_file(sim_paras, sim_par_names, t, path):
    """file_handling.py:20-45: names line, then the values separated by blanks."""
    with open(path + "sim_paras_t_" + _t(t) + ".txt", "w") as f:
        f.write(sim_par_names + "\n")
        for item in sim_paras:
            if isinstance(item, (list, np.ndarray)):
                for el in item:
                    f.write(f"{el} ")
            else:
                f.write(f"{item} ")


# ------------------------------------------------------------- particles
def save_particles_to_files(pos, cells, vel, m_w, m_s, xi, active_ids, vector_filename, scalar_filename,
                            cells_filename, xi_filename, active_ids_filename):
    """file_handling.py:47-99."""
    np.save(vector_filename, [pos, vel])
    np.save(cells_filename, cells)
    np.save(scalar_filename, [m_w, m_s])
    np.save(xi_filename, xi)
    np.save(active_ids_filename, active_ids)


def dump_particle_data(t, pos, vel, m_w, m_s, xi, T_grid, rv_grid, path):
    """file_handling.py:101-144: traced particles and the T / r_v grids at time t."""
    np.save(path + "particle_vector_data_" + _t(t) + ".npy", (pos, vel))
    np.save(path + "particle_scalar_data_" + _t(t) + ".npy", (m_w, m_s))
    np.save(path + "particle_xi_data_" + _t(t) + ".npy", xi)
    np.save(path + "grid_T_rv_" + _t(t) + ".npy", (T_grid, rv_grid))


def load_particle_data(path, save_times):
    """file_handling.py:146-186."""
    vec, scal, xi = [], [], []
    for t in save_times:
        vec.append(np.load(path + "particle_vector_data_" + _t(t) + ".npy"))
        scal.append(np.load(path + "particle_scalar_data_" + _t(t) + ".npy"))
        xi.append(np.load(path + "particle_xi_data_" + _t(t) + ".npy"))
    return np.array(vec), np.array(scal), np.array(xi)


def dump_particle_tracer_data_block(time_block, traced_vectors, traced_scalars, traced_xi, traced_water, path):
    """file_handling.py:188-232: one block of tracer dumps, named by its first time."""
    t = _t(time_block[0])
    np.save(path + "particle_vector_data_" + t + ".npy", traced_vectors)
    np.save(path + "particle_scalar_data_" + t + ".npy", traced_scalars)
    np.save(path + "particle_xi_data_" + t + ".npy", traced_xi)
    np.save(path + "water_removed_" + t + ".npy", traced_water)
    np.save(path + "particle_time_block_" + t + ".npy", time_block)


def load_particle_data_from_blocks(path, grid_save_times, pt_dumps_per_grid_frame):
    """file_handling.py:234-290: tracer blocks (4-D) unrolled, single dumps (3-D) appended."""
    vec, scal, xi = [], [], []
    for t in grid_save_times:
        v = np.load(path + "particle_vector_data_" + _t(t) + ".npy")
        s = np.load(path + "particle_scalar_data_" + _t(t) + ".npy")
        x = np.load(path + "particle_xi_data_" + _t(t) + ".npy")
        if v.ndim == 4:
            for n in range(pt_dumps_per_grid_frame):
                vec.append(v[n])
                scal.append(s[n])
                xi.append(x[n])
        elif v.ndim == 3:
            vec.append(v)
            scal.append(s)
            xi.append(x)
        else:
            raise ValueError(f"unexpected tracer array shape {v.shape}")
    return np.array(vec), np.array(scal), np.array(xi)


def dump_particle_data_all(t, pos, vel, cells, m_w, m_s, xi, active_ids, path):
    """file_handling.py:292-343: every particle at time t."""
    np.save(path + "particle_vector_data_all_" + _t(t) + ".npy", (pos, vel))
    np.save(path + "particle_cells_data_all_" + _t(t) + ".npy", cells)
    np.save(path + "particle_scalar_data_all_" + _t(t) + ".npy", (m_w, m_s))
    np.save(path + "particle_xi_data_all_" + _t(t) + ".npy", xi)
    np.save(path + "particle_active_ids_data_all_" + _t(t) + ".npy", active_ids)


def load_particle_data_all(path, save_times):
    """file_handling.py:345-412."""
    out = [[], [], [], [], []]
    names = ("particle_vector_data_all_", "particle_cells_data_all_", "particle_scalar_data_all_",
             "particle_xi_data_all_", "particle_active_ids_data_all_")
    for t in save_times:
        for k, n in enumerate(names):
            out[k].append(np.load(path + n + _t(t) + ".npy"))
    return tuple(np.array(a) for a in out)


# ------------------------------------------------------------------ grid
def save_grid_basics_to_textfile(grid, t, filename):
    """file_handling.py:415-434."""
    with open(filename, "w") as f:
        f.write("# grid.ranges[0] grid.ranges[1] grid.steps grid.step_y t\n")
        f.write(f"{grid.ranges[0][0]} {grid.ranges[0][1]} ")
        f.write(f"{grid.ranges[1][0]} {grid.ranges[1][1]} ")
        f.write(f"{grid.steps[0]} {grid.steps[1]} {grid.step_y} {t}")


def save_grid_arrays_to_npy_file(grid, filename1, filename2):
    """file_handling.py:436-464: arr_file1 = (p, T, rho_d, r_v, r_l, e_s, S, Theta), arr_file2 = (u, w, j_x, j_z)."""
    np.save(filename1, np.array([grid.pressure, grid.temperature, grid.mass_density_air_dry,
                                 grid.mixing_ratio_water_vapor, grid.mixing_ratio_water_liquid,
                                 grid.saturation_pressure, grid.saturation, grid.potential_temperature]))
    np.save(filename2, np.array([grid.velocity[0], grid.velocity[1], grid.mass_flux_air_dry[0],
                                 grid.mass_flux_air_dry[1]]))


def save_grid_scalar_fields(t, grid_scalar_fields, path, start_time=None):
    """file_handling.py:466-511: (r_v, r_l, Theta, T, p, S) at time t."""
    np.save(path + "grid_scalar_fields_t_" + _t(t) + ".npy",
            (grid_scalar_fields[4], grid_scalar_fields[5], grid_scalar_fields[2], grid_scalar_fields[0],
             grid_scalar_fields[1], grid_scalar_fields[6]))
    if start_time is not None:
        print("grid fields saved at t =", t, ", sim time:", datetime.now() - start_time)


def load_grid_scalar_fields(path, save_times):
    """file_handling.py:513-543."""
    return np.array([np.load(path + "grid_scalar_fields_t_" + _t(t) + ".npy") for t in save_times])


def save_grid_to_files(grid, t, basics_file, arr_file1, arr_file2):
    """file_handling.py:545-570."""
    save_grid_basics_to_textfile(grid, t, basics_file)
    save_grid_arrays_to_npy_file(grid, arr_file1, arr_file2)


def load_grid_from_files(basics_file, arr_file1, arr_file2):
    """file_handling.py:572-620: Grid with material properties and inverse dry densities recomputed."""
    basics = np.loadtxt(basics_file)
    scalars = np.load(arr_file1)
    vectors = np.load(arr_file2)
    grid = Grid([[basics[0], basics[1]], [basics[2], basics[3]]], [basics[4], basics[5]], basics[6])
    (grid.pressure, grid.temperature, grid.mass_density_air_dry, grid.mixing_ratio_water_vapor,
     grid.mixing_ratio_water_liquid, grid.saturation_pressure, grid.saturation,
     grid.potential_temperature) = [scalars[k] for k in range(8)]
    grid.update_material_properties()
    grid.rho_dry_inv = np.ones_like(grid.mass_density_air_dry) / grid.mass_density_air_dry
    grid.mass_dry_inv = (1.0 / grid.volume_cell) * grid.rho_dry_inv
    grid.velocity = np.array([vectors[0], vectors[1]])
    grid.mass_flux_air_dry = np.array([vectors[2], vectors[3]])
    return grid


def _full_names(t, path):
    t = _t(t)
    grid_files = [path + "grid_basics_" + t + ".txt", path + "arr_file1_" + t + ".npy",
                  path + "arr_file2_" + t + ".npy"]
    return (grid_files, path + "particle_vectors_" + t + ".npy", path + "particle_cells_" + t + ".npy",
            path + "particle_scalars_" + t + ".npy", path + "multiplicity_" + t + ".npy",
            path + "active_ids_" + t + ".npy")


def save_grid_and_particles_full(t, grid, pos, cells, vel, m_w, m_s, xi, active_ids, path):
    """file_handling.py:622-691: r_l from the particles (ascending id), then grid and particle files."""
    r_l = grid.mixing_ratio_water_liquid
    r_l.fill(0.0)
    act = np.arange(len(xi))[np.asarray(active_ids, dtype=bool)]
    np.add.at(r_l, (cells[0, act], cells[1, act]), m_w[act] * xi[act])
    r_l *= 1.0E-18 * grid.mass_dry_inv
    grid_files, vec_f, cells_f, scal_f, xi_f, act_f = _full_names(t, path)
    save_grid_to_files(grid, t, *grid_files)
    save_particles_to_files(pos, cells, vel, m_w, m_s, xi, active_ids, vec_f, scal_f, cells_f, xi_f, act_f)


def load_grid_and_particles_full(t, path):
    """file_handling.py:693-762: (grid, pos, cells, vel, m_w, m_s, xi, active_ids)."""
    grid_files, vec_f, cells_f, scal_f, xi_f, act_f = _full_names(t, path)
    grid = load_grid_from_files(*grid_files)
    vectors = np.load(vec_f)
    scalars = np.load(scal_f)
    return (grid, vectors[0], np.load(cells_f), vectors[1], scalars[0], scalars[1], np.load(xi_f),
            np.load(act_f))


def load_kernel_data_Ecol(kernel_method, save_dir_Ecol_grid, E_col_const=0.5):
    """file_handling.py:764-804: (E_col_grid, radius_grid, R_low, bin_factor, ln R_low, ln bin_factor, bins)."""
    if kernel_method == "Ecol_grid_R":
        radius_grid = np.load(os.path.join(save_dir_Ecol_grid, "radius_grid_out.npy"))
        E_col_grid = np.load(os.path.join(save_dir_Ecol_grid, "E_col_grid.npy"))
        R_kernel_low = radius_grid[0]
        bin_factor_R = radius_grid[1] / radius_grid[0]
        return (E_col_grid, radius_grid, R_kernel_low, bin_factor_R, np.log(R_kernel_low), np.log(bin_factor_R),
                len(radius_grid))
    if kernel_method == "Ecol_const":
        return E_col_const, None, None, None, None, None, None
    raise ValueError(f"unknown kernel_method {kernel_method!r} (Ecol_grid_R or Ecol_const)")
