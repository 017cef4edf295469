"""CPU oracle for the probebwa map path — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The product package probebwa_amd never
imports it.
"""
