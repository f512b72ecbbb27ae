"""Grid object of the 2-D kinematic case (drop-in for the reference's Grid).

Mirrors the attributes the SIP step reads and writes (grid.py:416-727 of the
reference): C-staggered Arakawa grid, cell fields (Nx, Nz), face fields
(2, Nx+1, Nz+1).  Only the numerics-relevant part is kept (no plotting).
Host-side only: set-up and frame I/O, not part of the time step.
"""
import math

import numpy as np

from .engine import host_material_properties


def compute_no_grid_cells_from_step_sizes(grid_ranges, grid_steps):
    """grid.py: number of cells so that the steps cover the ranges."""
    return [int(math.ceil((grid_ranges[0][1] - grid_ranges[0][0]) / grid_steps[0])),
            int(math.ceil((grid_ranges[1][1] - grid_ranges[1][0]) / grid_steps[1]))]


class Grid:
    """grid.py:416-727: geometry, thermodynamic fields and material properties."""

    def __init__(self, grid_ranges, grid_steps, dy):
        self.no_cells = np.array(compute_no_grid_cells_from_step_sizes(grid_ranges, grid_steps))
        self.no_cells_tot = self.no_cells[0] * self.no_cells[1]
        self.steps = np.array(grid_steps, dtype=np.float64)
        self.step_y = dy
        self.volume_cell = grid_steps[0] * grid_steps[1] * dy
        self.ranges = np.array(grid_ranges, dtype=np.float64)
        self.ranges[:, 1] = self.ranges[:, 0] + self.steps * self.no_cells
        self.sizes = np.array([self.ranges[0, 1] - self.ranges[0, 0], self.ranges[1, 1] - self.ranges[1, 0]])
        corners_x = np.linspace(0.0, self.sizes[0], self.no_cells[0] + 1) + self.ranges[0, 0]
        corners_y = np.linspace(0.0, self.sizes[1], self.no_cells[1] + 1) + self.ranges[1, 0]
        self.corners = np.array(np.meshgrid(corners_x, corners_y, indexing="ij"))
        self.centers = [self.corners[0][:-1, :-1] + 0.5 * self.steps[0],
                        self.corners[1][:-1, :-1] + 0.5 * self.steps[1]]
        shape = tuple(self.no_cells)
        for f in ("pressure", "temperature", "potential_temperature", "mass_density_air_dry",
                  "mixing_ratio_water_vapor", "mixing_ratio_water_liquid", "saturation_pressure", "saturation",
                  "heat_of_vaporization", "thermal_conductivity", "diffusion_constant", "surface_tension",
                  "specific_heat_capacity", "viscosity", "mass_density_fluid", "rho_dry_inv", "mass_dry_inv"):
            setattr(self, f, np.zeros(shape))
        self.velocity = np.zeros((2, self.no_cells[0] + 1, self.no_cells[1] + 1))
        self.mass_flux_air_dry = np.zeros_like(self.velocity)
        self.p_ref = 1.0E5
        self.p_ref_inv = 1.0E-5

    def update_material_properties(self):
        """grid.py:713-727 (host, libm)."""
        m = host_material_properties(self.temperature, self.pressure, self.mixing_ratio_water_vapor,
                                     self.mass_density_air_dry)
        (self.thermal_conductivity, self.diffusion_constant, self.heat_of_vaporization, self.surface_tension,
         self.specific_heat_capacity, self.viscosity, self.mass_density_fluid) = m

    def scalar_fields(self):
        """Stacked (10,Nx,Nz) array in the order of integration.py:1963-1972."""
        return np.array((self.temperature, self.pressure, self.potential_temperature, self.mass_density_air_dry,
                         self.mixing_ratio_water_vapor, self.mixing_ratio_water_liquid, self.saturation,
                         self.saturation_pressure, self.mass_dry_inv, self.rho_dry_inv))

    def mat_prop(self):
        """Stacked (7,Nx,Nz) array in the order of integration.py:1974-1980."""
        return np.array((self.thermal_conductivity, self.diffusion_constant, self.heat_of_vaporization,
                         self.surface_tension, self.specific_heat_capacity, self.viscosity,
                         self.mass_density_fluid))

    def set_from_stacked(self, scal, mat):
        (self.temperature, self.pressure, self.potential_temperature, self.mass_density_air_dry,
         self.mixing_ratio_water_vapor, self.mixing_ratio_water_liquid, self.saturation,
         self.saturation_pressure) = [np.array(a) for a in scal[:8]]
        (self.thermal_conductivity, self.diffusion_constant, self.heat_of_vaporization, self.surface_tension,
         self.specific_heat_capacity, self.viscosity, self.mass_density_fluid) = [np.array(a) for a in mat]
