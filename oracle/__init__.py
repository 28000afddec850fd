"""ORACLE — test infrastructure only.

A CPU restatement of the reference's contact phase in plain C++ (this folder), loaded
through ctypes. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
may import it, and only as the checker; the product never does.
"""
