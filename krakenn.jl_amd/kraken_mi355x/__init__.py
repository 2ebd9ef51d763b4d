"""kraken_mi355x — MI355X-native stream–collide backend for Kraken.jl's hot path.

The solver runs in hand-written CDNA4 HIP kernels (libkraken_mi355x.so, C ABI in
include/kraken_mi355x.h).  This package is the host-side mirror of the reference's
driver API; it never computes populations itself.
"""
from . import _lib, libb
from ._lib import KrakenError, load
from .batch import Batch, STATUS_NAMES, ensemble_parameters
from .drivers import (
    D2Q9,
    D3Q19,
    LBMConfig,
    cavity_3d_desc,
    cavity_desc,
    lbm_config,
    natural_convection_3d_desc,
    natural_convection_3d_initial_temperature,
    natural_convection_desc,
    natural_convection_initial_temperature,
    omega,
    rayleigh_benard_desc,
    rb_initial_temperature,
    reynolds,
    run_cavity_2d,
    run_cavity_3d,
    run_couette_2d,
    run_natural_convection_2d,
    run_natural_convection_3d,
    run_poiseuille_2d,
    run_rayleigh_benard_2d,
    run_taylor_green_2d,
    run_taylor_green_3d,
    run_to_residual,
    taylor_green_2d_fields,
    taylor_green_3d_desc,
    taylor_green_3d_fields,
    trt_rates,
)
from .lattice import ArgumentError, Lattice, MI355XBackend, make_desc
from .libb import (cut_links_from_dense, precompute_q_wall_cylinder, precompute_q_wall_sphere_3d,
                   run_cylinder_libb_2d, run_sphere_libb_3d,
                   wall_velocity_rotating_cylinder)
from .runner import (BoundarySetup, DomainSetup, GeometryRegion, InitialSetup, PhysicsSetup,
                     SimulationSetup, plan_simulation, run_simulation)

__all__ = [
    "BoundarySetup", "DomainSetup", "GeometryRegion", "InitialSetup", "PhysicsSetup", "SimulationSetup",
    "plan_simulation", "run_simulation", "run_cylinder_libb_2d", "run_sphere_libb_3d", "cut_links_from_dense", "precompute_q_wall_cylinder",
    "precompute_q_wall_sphere_3d", "wall_velocity_rotating_cylinder",
    "Batch", "STATUS_NAMES", "ensemble_parameters", "ArgumentError", "D2Q9", "D3Q19", "KrakenError", "LBMConfig", "Lattice", "MI355XBackend",
    "cavity_3d_desc", "cavity_desc", "lbm_config", "load", "make_desc", "natural_convection_3d_desc", "natural_convection_3d_initial_temperature", "natural_convection_desc",
    "natural_convection_initial_temperature", "omega", "rayleigh_benard_desc",
    "rb_initial_temperature", "reynolds", "run_cavity_2d", "run_cavity_3d", "run_couette_2d", "run_natural_convection_2d", "run_natural_convection_3d",
    "run_poiseuille_2d", "run_rayleigh_benard_2d", "run_taylor_green_2d", "run_taylor_green_3d", "run_to_residual", "taylor_green_3d_desc",
    "taylor_green_2d_fields", "taylor_green_3d_fields", "trt_rates",
]
