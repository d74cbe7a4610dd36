"""The device library loads on a CPU-only host and exports every symbol its
header declares (no device calls)."""
import ctypes

import capi
import capi_m8


def test_library_exports_declared_symbols():
    names = capi.declared_symbols()
    assert len(names) >= 14
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_sizes_match_header_layout():
    # Field order/padding of the ctypes mirrors (x86-64 SysV layout).
    assert ctypes.sizeof(capi.IndexDesc) == 4 * 8 + 4 + 4 + 8 + 7 * 8
    assert ctypes.sizeof(capi.IndexArrays) == 6 * 8
    assert ctypes.sizeof(capi.SearchArgs) == 8 * 4 + 8 + 4 * 5 + 4 + 8 + 8 + 8
    assert ctypes.sizeof(capi.CtxInfo) == 4 * 4 + 3 * 8


def test_me8q_struct_sizes():
    assert ctypes.sizeof(capi_m8.M8Desc) == 4 * 8 + 8 + 6 * 8 + 8
    assert ctypes.sizeof(capi_m8.M8Arrays) == 3 * 8
    # ... outputs_on_device, mode, batched_budget, batched_set_a
    assert ctypes.sizeof(capi_m8.M8SearchArgs) == 5 * 8 + 4 + 4 + 8 + 8 + 4 + 4 + 4 + 4
    # ... groups, form (+ padding), overflowed, pairs_scored, items
    assert ctypes.sizeof(capi_m8.M8Stats) == 5 * 8 + 2 * 8 + 3 * 8 + 8 + 8 + 8 + 8 + 3 * 2 * 8
