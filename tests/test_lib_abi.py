"""CPU checks of the C-ABI library: it loads and exports every declared symbol."""
import ctypes
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols(header: str = "felis.h"):
    text = (ROOT / "include" / header).read_text()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(felis_\w+)\s*\(", text, flags=re.M)))


def test_header_matches_binding_list():
    from felis_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_every_symbol():
    from felis_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    for name in declared_symbols():
        assert re.search(rf"\bT {name}\b", nm), name


def test_hdf_library_exports_every_symbol():
    """libfelis_hdf.so (HDF5 dump/restore) loads next to the core library and
    exports every entry point include/felis_hdf.h declares."""
    from felis_amd import _lib
    assert sorted(_lib.HDF_EXPORTS) == declared_symbols("felis_hdf.h")
    lib = _lib.load_hdf()
    for name in declared_symbols("felis_hdf.h"):
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", str(_lib.HDF_PATH)], capture_output=True, text=True).stdout
    for name in declared_symbols("felis_hdf.h"):
        assert re.search(rf"\bT {name}\b", nm), name


def test_create_without_gpu_fails_cleanly():
    """No compute without a GPU: felis_create reports a HIP error instead of crashing."""
    import torch
    if torch.cuda.is_available():
        return
    from felis_amd import _lib
    import fel_cases as fc
    d = fc.make_desc(fc.make_case())
    h = ctypes.c_void_p()
    rc = _lib.load().felis_create(ctypes.byref(d), ctypes.byref(h))
    assert rc != 0
