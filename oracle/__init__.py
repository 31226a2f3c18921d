"""CPU FP64 oracle of the radiation-field / level-population path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product (magritte_amd) never imports it.

Parity status of each oracle function is listed in the module docstring of
oracle/py_oracle.py.
"""
