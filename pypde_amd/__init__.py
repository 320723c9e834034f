"""MI355X-native ADER-WENO solver with the call surface of haranjackson/PyPDE.

    from pypde_amd import pde_solver, weno_solver, systems
"""
from pypde_amd.solvers import Solver, pde_solver, weno_solver  # noqa: F401
from pypde_amd import systems  # noqa: F401
