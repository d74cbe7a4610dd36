"""Argument validation of the e8qlib.E8qIndex Python surface (no device work):
same exception types and messages as the reference binding
(python_bindings/e8q_bindings.cpp:24-99, 237-252)."""
import numpy as np
import pytest

import e8qlib


def test_constructor_validation():
    with pytest.raises(ValueError, match="dim must be positive"):
        e8qlib.E8qIndex(0, 4)
    with pytest.raises(ValueError, match=r"nbits must be in \[1, 13\]"):
        e8qlib.E8qIndex(16, 0)
    with pytest.raises(ValueError, match=r"nbits must be in \[1, 13\]"):
        e8qlib.E8qIndex(16, 14)
    ix = e8qlib.E8qIndex(dim=16, nbits=4, seed=3)
    assert ix.dim == 16 and ix.nbits == 4 and not ix.is_built and ix.n == 0


def test_unbuilt_errors():
    ix = e8qlib.E8qIndex(16, 4)
    q = np.zeros((2, 16), np.float32)
    with pytest.raises(RuntimeError, match="must be built or loaded before search"):
        ix.search(q, 10, 4)
    with pytest.raises(RuntimeError, match="before save"):
        ix.save("/tmp/never.e8ax")
    with pytest.raises(RuntimeError, match="before set_prune_alpha"):
        ix.set_prune_alpha(0.5)
    with pytest.raises(RuntimeError, match="before set_disable_prune"):
        ix.set_disable_prune(True)


def test_build_shape_validation():
    ix = e8qlib.E8qIndex(16, 4)
    data = np.zeros((10, 16), np.float32)
    cent = np.zeros((2, 16), np.float32)
    ids = np.zeros(10, np.uint32)
    with pytest.raises(ValueError, match="data dimension does not match"):
        ix.build(np.zeros((10, 8), np.float32), cent, ids)
    with pytest.raises(ValueError, match="centroid dimension does not match"):
        ix.build(data, np.zeros((2, 8), np.float32), ids)
    with pytest.raises(ValueError, match="cluster_ids length must match"):
        ix.build(data, cent, np.zeros(9, np.uint32))
    with pytest.raises(ValueError, match="must be a 2D NumPy array"):
        ix.build(np.zeros(16, np.float32), cent, ids)


def test_load_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.e8ax"
    p.write_bytes(b"\x00" * 64)
    with pytest.raises(RuntimeError, match="not an E8Q index file"):
        e8qlib.E8qIndex.load(str(p))
    with pytest.raises(RuntimeError, match="cannot open index file"):
        e8qlib.E8qIndex.load(str(tmp_path / "missing.e8ax"))


def test_extension_surface():
    # Search forms, launch options and serving beside the reference's surface.
    for name in ("set_search_graphs", "set_batched_shape", "search_context_info", "Server"):
        assert hasattr(e8qlib, name), name
    ix = e8qlib.E8qIndex(16, 4)
    for name in ("search_device", "build_device"):
        assert hasattr(ix, name), name
    with pytest.raises(ValueError, match="null device pointer"):
        ix.build_device(0, 0, 2, np.zeros(4, np.uint32), 0)
