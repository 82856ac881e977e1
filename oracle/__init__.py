"""CPU oracles (test infrastructure only): ntoracle.c (propagation) and
rqflow.py (spline flow)."""
