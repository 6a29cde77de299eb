"""CPU-only checks of the drop-in boundary: the library loads without a GPU,
exports every entry point include/octohip.h declares, and the front-end's
argument checks raise the reference's error types before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import octohip
from octohip import _lib, datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("octohip.h",):
        src = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(oct_\w+)\s*\(", src, re.M))
    return names


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("oct_create", "oct_run", "oct_build_edges", "oct_build_edges_backward", "oct_device_view"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_covers_every_declared_symbol():
    assert declared_functions() == set(_lib.SIGNATURES)
    _lib.load()


def test_struct_sizes_match_header():
    # field layouts the kernels and the header agree on
    assert ctypes.sizeof(_lib.OctEdges) == 40
    assert ctypes.sizeof(_lib.OctTree) == 6 * 8 + 4 + 4 + 8 + 12 + 4
    assert ctypes.sizeof(_lib.OctView) == 10 * 8 + 4 * 4 + 8 + 8
    assert ctypes.sizeof(_lib.OctEdgeStats) == 3 * 8 + 4 * 8 + 8 + 4 * 4
    assert ctypes.sizeof(_lib.OctBuildInfo) == 16
    assert ctypes.sizeof(_lib.OctNeighbors) == 2 * 8 + 2 * 8 + 2 * 4
    assert ctypes.sizeof(_lib.OctBatchEdges) == 4 * 8 + 8 + 8 + 2 * 4 + 3 * 8


def test_argument_errors_without_device():
    tns = octohip.TNS()
    with pytest.raises(octohip.ArgumentError):
        octohip.set_search_radius(tns, 0.0)
    with pytest.raises(octohip.ArgumentError):
        octohip.set_search_radius(tns, -1.0)
    with pytest.raises(octohip.ArgumentError):
        octohip.run(tns)  # no radius
    octohip.set_search_radius(tns, 0.1)
    with pytest.raises(octohip.ArgumentError):
        octohip.run(tns)  # no point sets
    with pytest.raises(octohip.ArgumentError):
        octohip.TNS(ndims=4)
    with pytest.raises(octohip.OctopusError):
        octohip.for_each_neighbor(lambda j: None, tns, 0, 0, 0)


def test_coords_validation_without_device():
    import torch

    tns = octohip.TNS()
    with pytest.raises(octohip.DimensionMismatch):
        octohip.add_point_set(tns, torch.zeros(10, 2))
    with pytest.raises(octohip.ArgumentError):
        octohip.add_point_set(tns, torch.zeros(10, 3))  # CPU tensor on a device engine


def test_datasets_deterministic():
    a = datasets.uniform_cube(1000, 3)
    b = datasets.uniform_cube(1000, 3)
    assert a.dtype == np.float32 and np.array_equal(a, b)
    assert a.min() >= 0 and a.max() < 1
    assert not np.array_equal(a, datasets.uniform_cube(1000, 4))
    z = datasets.normal(1, 0, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_sph_lattice_shape():
    pts, r = datasets.sph_lattice(side=10)
    assert pts.shape == (1000, 3) and np.float32(r) == np.float32(0.02)
    g = np.round(pts / 0.01)
    assert np.all(np.abs(pts - g * 0.01) <= 0.001 + 1e-7)


def test_device_views_hold_their_owner():
    """A wrapped result keeps the TNS (or batch) whose context owns its
    memory alive: the DLPack struct torch holds for the tensor keeps the
    owner referenced until torch's deleter releases it."""
    import ctypes
    import gc
    import weakref

    class Owner:
        pass

    owner = Owner()
    ref = weakref.ref(owner)
    cap = octohip._dl_capsule(0x1000, (4, 3), "<f4", 0, owner)
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    addr = get(cap, b"dltensor")
    dm = octohip._DLManagedTensor.from_address(addr)
    assert dm.dl_tensor.data == 0x1000 and dm.dl_tensor.ndim == 2
    assert (dm.dl_tensor.device.device_type, dm.dl_tensor.device.device_id) == (10, 0)
    assert (dm.dl_tensor.dtype.code, dm.dl_tensor.dtype.bits, dm.dl_tensor.dtype.lanes) == (2, 32, 1)
    assert [dm.dl_tensor.shape[i] for i in range(2)] == [4, 3] and not dm.dl_tensor.strides
    del owner
    gc.collect()
    assert ref() is not None  # held by the handed-out struct
    dm.deleter(ctypes.pointer(dm))  # what torch calls when the tensor dies
    gc.collect()
    assert ref() is None


def test_views_held_at_interpreter_exit():
    """Result tensors a script still holds when the interpreter exits are
    freed after the module's own dict is cleared: their deleter and the
    structs torch reads must still exist then (the process used to end with
    a segmentation fault).  Host memory stands in for the device buffers
    (kDLCPU), so this runs without a GPU."""
    import subprocess
    import sys

    code = (
        "import sys; sys.path[:0] = [{pkg!r}]\n"
        "import numpy as np, torch, octohip\n"
        "octohip._KDL_ROCM = 1  # kDLCPU: host memory in place of device buffers\n"
        "buf = np.arange(64, dtype=np.float32)\n"
        "held = [torch.from_dlpack(octohip._dl_capsule(buf.ctypes.data, (64,), '<f4', 0, buf)) for _ in range(4)]\n"
        "assert float(held[3][5]) == 5.0\n"
    ).format(pkg=os.path.join(ROOT, "octopus.jl_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])


def test_result_views_are_slices_of_kept_views(monkeypatch):
    """build_edges returns slices of a view kept per (pair, field): the same
    address and a length within the kept view give a slice of it; a longer
    length or a new address (a grown buffer) give a new view.  Host memory
    stands in for the device buffers (kDLCPU)."""
    import torch

    monkeypatch.setattr(octohip, "_KDL_ROCM", 1)
    tns = octohip.TNS()
    tns.device = torch.device("cpu", 0)
    a = np.arange(400, dtype=np.float32)
    b = np.arange(400, dtype=np.float32) + 1000
    pa, pb = a.ctypes.data, b.ctypes.data
    v1 = tns._view((0, 0, 2), pa, 100, (3,), torch.float32)
    assert v1.shape == (100, 3) and v1.data_ptr() == pa and float(v1[1, 0]) == 3.0
    base = tns._views[(0, 0, 2)][1]
    v2 = tns._view((0, 0, 2), pa, 60, (3,), torch.float32)
    assert v2.shape == (60, 3) and tns._views[(0, 0, 2)][1] is base  # a slice of the kept view
    v3 = tns._view((0, 0, 2), pa, 130, (3,), torch.float32)
    assert v3.shape == (130, 3) and tns._views[(0, 0, 2)][1] is not base  # longer: wrapped again
    v4 = tns._view((0, 0, 2), pb, 10, (3,), torch.float32)
    assert v4.data_ptr() == pb and float(v4[0, 0]) == 1000.0  # a new address: wrapped again
    assert tns._view((0, 0, 0), pa, 0, (), torch.int32).shape == (0,)
    assert float(v1[2, 2]) == 8.0  # earlier results still view the first buffer
