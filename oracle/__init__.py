"""TEST INFRASTRUCTURE ONLY: the CPU oracle for GITM's vertical step.

``oracle.cpu`` wraps the C restatement (liboracle.so, built by
``oracle/Makefile``); ``oracle.ref`` drives the reference Fortran build in
``oracle/_ref`` when it exists.  Only tests/, the smoke check and bench.py's
CPU-baseline leg use this package.
"""
