"""tictac: MI355X-native AGS/WPCD solve path of Tic-tac (gfx950 HIP behind a C-ABI).

Python side of the library: a ctypes binding of include/tictac.h and a thin
mirror of the reference driver.  All numerics run in libtictac.so.
"""
from ._lib import (SOLVE_DENSE, SOLVE_GMRES, SOLVE_MIXED, SOLVE_PADE, TictacError, cell_tuning, lib)  # noqa: F401
from .problem import (Channel, Context, P123, Potential, Problem, chebyshev_scale, run_params)  # noqa: F401
from .synthetic import synthetic_params  # noqa: F401
