"""The C-ABI library loads and exports every entry point include/nddo.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nddo.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nddo_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("nddo_engine_create", "nddo_engine_destroy", "nddo_scf_once", "nddo_scf_batch", "nddo_submit",
                 "nddo_poll", "nddo_wait", "nddo_get_result", "nddo_release", "nddo_last_error",
                 "nddo_dbg_integrals", "nddo_dbg_fock", "nddo_dbg_eigh", "nddo_dbg_density"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from molds_amd.nddo import LIBPATH
    assert os.path.exists(LIBPATH), "build libnddo.so first (python -m molds_amd.build)"
    lib = ctypes.CDLL(LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_opts_default_matches_reference_parameters():
    # src/base/Parameters.cpp:132-144 (host-only call)
    from molds_amd.nddo import default_opts
    o = default_opts()
    assert (o.max_iter, o.rms_density, o.damping_thresh, o.damping_weight, o.diis_num_error_vect,
            o.diis_start_error, o.diis_end_error) == (100, 1e-8, 1.0, 0.8, 5, 1e-2, 1e-8)


def test_engine_without_gpu_fails_loudly():
    """No CPU fallback: with no HIP device the engine refuses to start."""
    import pytest
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    from molds_amd.nddo import Engine, NddoError
    with pytest.raises(NddoError):
        Engine(0)


def test_stage_names_match_header_enum():
    """The binding's stage list has one name per NDDO_STAGE_* entry, in order."""
    from molds_amd.nddo import Engine
    txt = open(HEADER).read()
    body = re.search(r"enum\s*\{\s*(NDDO_STAGE_SETUP[^}]*)\}", txt).group(1)
    names = [t.strip().split("=")[0].strip() for t in body.split(",") if t.strip()]
    assert names[-1] == "NDDO_STAGE_COUNT"
    assert len(Engine.STAGES) == len(names) - 1
    assert [n[len("NDDO_STAGE_"):].lower() for n in names[:-1]][:6] == list(Engine.STAGES[:6])


def test_result_struct_matches_header():
    """ctypes nddo_result fields follow the header's declaration order."""
    from molds_amd.nddo import Result
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\} nddo_result;", txt).group(1)
    decl = [re.findall(r"([A-Za-z_0-9]+)\s*;", line)[0] for line in body.split("\n") if ";" in line]
    assert [f[0] for f in Result._fields_] == decl
