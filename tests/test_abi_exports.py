"""CPU checks of the C-ABI library: it loads, and it exports every function
that include/*.h declares (no device calls are made here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "zkp-dynaverse_amd", "dvgpu", "libdynaverse_hip.so")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(?:int|void)\s+(dvgpu_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from dvgpu import _lib
    assert declared_functions() <= set(_lib.SIGNATURES)


def test_struct_layouts_match_the_reference_abi():
    from dvgpu import _lib
    sizes = {
        _lib.DvgpuResidentOp: 104,            # dynaverse_gpu.h:119-125, resident.rs:41-47
        _lib.DvgpuG2ValidationReport: 32,     # icicle_g2_validation.cu:14
        _lib.DvgpuG2ValidationStats: 64,
        _lib.DvgpuG2KernelConfig: 16,         # icicle_g2_validation.cu:15
        _lib.DvgpuG2KernelResources: 56,
        _lib.DvgpuG2ExperimentStats: 24,
        _lib.DvgpuBatchStats: 32,
        _lib.DvgpuSetupStats: 13 * 8,
    }
    for cls, size in sizes.items():
        assert ctypes.sizeof(cls) == size, cls.__name__
