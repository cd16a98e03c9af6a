"""fvmhip: MI355X (gfx950) hot path of XCALibre_TW.jl behind its own API.

    from fvmhip import *
    mesh = build_mesh("quad40.unv", scale=0.001)
    model = RANS(Laminar, mesh=mesh, viscosity=ConstantScalar(0.01))
    assign(model, "U", Dirichlet("inlet", [0, 0, 0]), ...)
    ...
    Rx, Ry, Rp = simple(model, config, pref=0.0)
"""
from ._lib import FVHError, device_count, lib  # noqa: F401
from .api import (  # noqa: F401
    DILU, ILU0, LDL, BicgstabSolver, CgSolver, Configuration, ConstantScalar, Dirichlet, Euler, GmresSolver, Jacobi,
    KOmega, KWallFunction, Laminar, Linear, Midpoint, Neumann, NormDiagonal, NutWallFunction, OmegaWallFunction,
    Orthogonal, RANS, BatchSimulation, CaseDesc, ScalarEquation, Simulation, Steady, Upwind, assign, initialise,
    isturbulent, model2vtk, piso, pressure_force, set_runtime, set_schemes, set_solver, simple, surface_flux,
    taylor_green_flux, viscous_force, wall_shear_stress, write_vtk)
from .mesh import (Mesh2, backward_facing_step, build_mesh, cavity, channel, delaunay_unit_square,  # noqa: F401
                   flat_plate, from_elements, rect_grid, union)
