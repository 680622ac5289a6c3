"""Test infrastructure: the CPU oracle for the DirectPseudoSpectral solver.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product (mpi_amd) never does.
"""
