"""ORACLE -- test infrastructure only (CPU FP64 restatement of the hot path).

Importable only from tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  Product code never imports it.
"""
