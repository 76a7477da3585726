"""CPU oracle for the BN254 hot path -- test infrastructure only.

Importable from tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg; the product package (zkp-dynaverse_amd/dvgpu) never imports it.
"""
