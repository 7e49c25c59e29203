"""``import aspher`` — the reference module name, served by aspher_amd."""
from aspher_amd.api import (ContactResult, ContactSolver, configure_allocator,  # noqa: F401
                            solve_nested)
