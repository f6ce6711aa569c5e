"""FP64 CPU oracle for the three-phonon scattering + thermal-conductivity path.

TEST INFRASTRUCTURE ONLY. This package restates, in plain numpy (and a small C
file for the per-element reference loop), the algorithm of TDEP's
`thermal_conductivity` program.  It is imported only by `tests/`, by
`__graft_entry__.smoke()` (as the checker) and by the `cpu_baseline` leg of
`bench.py`.  The product path (`tdep_amd`) never imports, links or executes
anything in this directory.
"""
