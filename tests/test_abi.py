"""The C-ABI libraries load and export every entry point include/*.h
declares (CPU only: no device calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "physis", "physis_hip.h")
LIB = os.path.join(ROOT, "physis_amd", "lib")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"extern\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_runtime_abi():
    names = declared_functions()
    for must in ("PSInit", "PSFinalize", "__PSGridNew", "__PSGridFree",
                 "__PSGridCopyin", "__PSGridCopyout", "__PSReduceGridFloat",
                 "__PSReduceGridDouble", "__PSReduceGridInt",
                 "__PSReduceGridLong", "__PSGridSet", "__PSGridGetID",
                 "__PSGridSwap", "__PSCheckHipError"):
        assert must in names


def test_runtime_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(LIB, "libphysis_rt_hip.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    for n in ("PSDomain1DNew", "PSDomain2DNew", "PSDomain3DNew"):
        assert hasattr(lib, n)
    assert not missing, missing


@pytest.mark.parametrize("prog,funcs", [
    ("diffusion7", ["diffusion7_run", "diffusion7_periodic_run",
                    "diffusion7_sweep_domain"]),
    ("himeno", ["himeno_run", "himeno_run_gosa"])])
def test_programs_export_their_functions(prog, funcs):
    C.CDLL(os.path.join(LIB, "libphysis_rt_hip.so"), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(os.path.join(LIB, f"libphysis_prog_{prog}.so"))
    for f in funcs:
        assert hasattr(lib, f)


def test_poisson_library_exports_its_header():
    text = open(os.path.join(ROOT, "include", "physis",
                             "physis_poisson.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(PSPoisson\w+)\s*\(", text)))
    assert "PSPoissonSolveMixed" in names and "PSPoissonNew" in names
    C.CDLL(os.path.join(LIB, "libphysis_rt_hip.so"), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(os.path.join(LIB, "libphysis_poisson_hip.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_warmstart_library_exports_its_header():
    text = open(os.path.join(ROOT, "include", "physis",
                             "physis_warmstart.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(PS(?:UNet|Conv3d)\w+)\s*\(", text)))
    assert "PSUNetTrainStep" in names and "PSConv3dForwardHost" in names
    C.CDLL(os.path.join(LIB, "libphysis_rt_hip.so"), mode=C.RTLD_GLOBAL)
    lib = C.CDLL(os.path.join(LIB, "libphysis_warmstart_hip.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
