"""Physis HIP target for MI355X: libphysis_rt_hip, the translated programs
and their Python host mirror.  See runtime.py (grids, reductions) and
programs.py (the benchmark programs' functions)."""
