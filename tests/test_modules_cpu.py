"""MuhRec plugin boundary on CPU: the plugin quartet, module lists and the
error behaviour the reference's loader and modules show before any device
work (ModuleItem.cpp:93-155, ParameterHandling.cpp:25-37,
ProjectionFilter.cpp:265-266, ReconConfig checks)."""
import ctypes

import pytest

from imagingsuite_amd import engine


BP_PARAMS = {"ProjectionBufferSize": 16, "SliceBlock": 32, "SubVolume": "1 1"}
FILTER_PARAMS = {"filtertype": "Hamming", "cutoff": 0.5, "order": 1, "usebias": "true",
                 "biasweight": 0.1, "paddingdoubler": 2}
CFG = {"geometry": "parallel", "center": 32, "roi": "0 0 64 16", "resolution": "0.1 0.1"}


def test_plugin_quartet():
    so = ctypes.CDLL(engine.MODULES_PATH)
    so.GetModule.restype = ctypes.c_void_p
    so.GetModule.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p]
    so.Destroy.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    assert so.LibVersion() == -1
    assert so.GetModule(b"kiptool", b"MultiProjBPparallel", None) is None
    assert so.GetModule(b"muhrecbp", b"NoSuchModule", None) is None
    assert so.Destroy(b"kiptool", None) == -1
    m = so.GetModule(b"muhrecbp", b"MultiProjBPparallel", None)
    assert m
    assert so.Destroy(b"muhrecbp", ctypes.c_void_p(m)) == 0


def test_module_lists():
    bp = engine.module_list("muhrecbp")
    for name in ["MultiProjBPparallel", "MultiProjBP", "NearestNeighborBP", "FDKbp_single", "FDKbp"]:
        assert name in bp
        assert bp[name]["ProjectionBufferSize"] and bp[name]["SliceBlock"]
    assert bp["MultiProjBPparallel"]["ProjectionBufferSize"] == "16"
    assert bp["FDKbp_single"]["ProjectionBufferSize"] == "2000"
    pre = engine.module_list("muhrec")
    assert set(pre["ProjectionFilterSingle"]) == {"filtertype", "cutoff", "order", "usebias",
                                                  "biasweight", "paddingdoubler"}
    assert pre["ProjectionFilterSingle"]["filtertype"] == "Hamming"
    assert set(pre["FullLogNorm"]) == {"usenormregion", "uselut", "referenceaverage"}
    assert pre["FullLogNorm"]["uselut"] == "false"
    assert engine.module_list("other")["__rc__"] == -1


def test_missing_parameter_is_module_exception():
    p = dict(BP_PARAMS)
    del p["SliceBlock"]
    with pytest.raises(engine.EngineError, match="Could not find parameter SliceBlock"):
        engine.Engine("MultiProjBPparallel", p, CFG)


def test_cutoff_out_of_range():
    f = dict(FILTER_PARAMS, cutoff=0.7)
    with pytest.raises(engine.EngineError, match="Cut off frequency"):
        engine.Engine("MultiProjBPparallel", BP_PARAMS, CFG, "ProjectionFilterSingle", f)


def test_unknown_filter_and_module():
    f = dict(FILTER_PARAMS, filtertype="Gauss")
    with pytest.raises(engine.EngineError, match="filter string"):
        engine.Engine("MultiProjBPparallel", BP_PARAMS, CFG, "ProjectionFilterSingle", f)
    with pytest.raises(engine.EngineError, match="Failed to create module"):
        engine.Engine("NoSuchBP", BP_PARAMS, CFG)


def test_missing_library():
    with pytest.raises(engine.EngineError, match="Could not load module library"):
        engine.Engine("MultiProjBPparallel", BP_PARAMS, CFG, module_so="/nonexistent/lib.so")


def test_configure_valid_without_device():
    e = engine.Engine("MultiProjBPparallel", BP_PARAMS, CFG, "ProjectionFilterSingle", FILTER_PARAMS)
    e.close()


def test_modules_header_symbols_exported():
    """Every function include/muhrec_modules.h declares is exported: the plugin
    quartet and mr_modules_* by the module library, mre_* by the driver."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    txt = open(os.path.join(root, "include", "muhrec_modules.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", txt, re.M))
    assert {"GetModule", "Destroy", "LibVersion", "GetModuleList", "mr_modules_device_handoffs",
            "mre_engine_create", "mre_engine_run", "mre_engine_phase_times"} <= names
    mods, eng = ctypes.CDLL(engine.MODULES_PATH), ctypes.CDLL(engine.ENGINE_PATH)
    for n in sorted(names):
        lib = eng if n.startswith("mre_") else mods
        assert hasattr(lib, n), n
    assert engine.device_handoffs() >= 0


def test_abort_after_options_parse_without_device():
    """bp_abort_after / filter_abort_after give the back-projector and
    ProjectionFilterSingle an interactor; a value that is not an integer is
    refused when the engine is made."""
    for key in ("bp_abort_after", "filter_abort_after"):
        e = engine.Engine("MultiProjBPparallel", BP_PARAMS, dict(CFG, **{key: 2}), "ProjectionFilterSingle",
                          FILTER_PARAMS)
        e.close()
        with pytest.raises(engine.EngineError):
            engine.Engine("MultiProjBPparallel", BP_PARAMS, dict(CFG, **{key: "soon"}), "ProjectionFilterSingle",
                          FILTER_PARAMS)
