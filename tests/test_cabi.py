"""CPU checks of the drop-in boundary: the built library loads and exports
every entry point include/csm_c.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "csm_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csm_[a-z_]+)\s*\(", text)))


def test_header_declares_the_reference_surface():
    names = _declared()
    for must in ("csm_create", "csm_preprocess", "csm_initial", "csm_classify", "csm_add_edge",
                 "csm_remove_edge", "csm_add_vertex", "csm_remove_vertex", "csm_apply_batch",
                 "csm_totals", "csm_memory", "csm_last_error", "csm_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from paracosm_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("engine library not built (python -m paracosm_amd.build)")
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(engine.exported_symbols()) == _declared()


def test_python_mirror_refuses_to_run_without_library(monkeypatch, tmp_path):
    from paracosm_amd import engine
    monkeypatch.setattr(engine, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(engine, "_lib", None)
    with pytest.raises(ImportError):
        engine.lib()


def test_update_record_layout_matches_insertunit():
    """csm_update mirrors InsertUnit (utils/types.h:34-42) as 16 packed bytes."""
    from paracosm_amd.engine import UPDATE_DTYPE, make_updates
    assert UPDATE_DTYPE.itemsize == 16
    u = make_updates([1, 0], [3, 4], [5, 6], [7, 0])
    assert bytes(u[0].tobytes()[:2]) == b"e\x01"
    assert u["id1"].tolist() == [3, 4] and u["id2"].tolist() == [5, 6]
