"""TEST INFRASTRUCTURE ONLY: the CPU oracle for the gridding path.

Importable only from tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, and only as the checker; the product never imports it.
"""
