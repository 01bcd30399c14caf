"""CPU oracle of the explicit step — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package, and only as the checker (or the timed CPU baseline), never as a
product path.  See q2d_oracle.c for the reference files it restates.
"""
from .binding import (Oracle, build, cons2prim, inflow_flux, lib_path, load, outflow_flux, planted,  # noqa: F401
                      pow_, prim2cons, thermo, unit_tables)
