"""MI355X-native BSSN evolution: drop-in for the reference's RHS + RK4 path.

Public names mirror the reference (uwplasma/JAX_BSSN):
``BSSNVariables``, ``BSSNParameters`` (bssn/variables.py),
``compute_bssn_rhs``, ``enforce_algebraic_constraints``, ``rk4_step``
(evolution/time_evolve.py), ``compute_all_constraints``
(bssn/constraints.py), ``compute_constraint_norms``
(diagnostics/constraints.py).  All compute runs in hand-written HIP kernels
for gfx950 behind the C ABI of include/bssn_amd.h.
"""

from .variables import BSSNParameters, BSSNVariables, get_boundary_codes, pack, unpack  # noqa: F401
from .time_evolve import (  # noqa: F401
    Evolver,
    compute_bssn_rhs,
    compute_bssn_rhs_with_matter,
    enforce_algebraic_constraints,
    evolve,
    evolve_packed,
    rhs_packed,
    rk4_step,
)
from .constraints import (  # noqa: F401
    ConstraintViolations,
    compute_all_constraints,
    compute_all_constraints_with_matter,
    compute_constraint_norms,
    compute_hamiltonian_constraint,
)

__version__ = "0.1.0"
