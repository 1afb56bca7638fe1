"""CPU oracle of the PMM likelihood / EM path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker; the product (laml_amd) never
imports it.
"""
