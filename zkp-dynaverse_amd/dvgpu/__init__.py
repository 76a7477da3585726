"""MI355X-native BN254 proving hot path for ZKP-Dynaverse.

``HipBackend`` is the operator backend (the reference's ``IcicleBackend``
counterpart) over ``libdynaverse_hip.so``; ``codec`` is the wire format and
``rng`` the seeded ChaCha20 workload generator.
"""
from .backend import BackendError, BatchMsmRequest, HipBackend, SparseMsmRequest  # noqa: F401
from . import codec, rng  # noqa: F401
