"""CPU restatement oracle of the SSE chain -- test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker of the GPU path; the product never imports it.
"""
