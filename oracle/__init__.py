"""The oracle: a CPU restatement of the reference's serial simulation hot
path (oracle.c + stepper.py) and the recipe that builds the reference's own
serial library (Makefile.ref, into _ref/).

TEST INFRASTRUCTURE ONLY: it is what the GPU path is checked against.  Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
anything here.
"""
