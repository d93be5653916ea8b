"""The C-ABI library builds, loads and exports every entry point of
include/bass_hip.h; host-side argument validation works without a GPU."""

import ctypes
import os
import re

import pytest

from conftest import ROOT, pkg_module

HEADER = os.path.join(ROOT, "include", "bass_hip.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bass_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    L = pkg_module("_lib")
    if not os.path.exists(L.LIB_PATH):
        import __graft_entry__

        __graft_entry__.build()
    return L.load()


def test_header_declares_expected_surface():
    names = header_functions()
    for must in ["bass_gate", "bass_truncate", "bass_diag2q", "bass_rdm1q", "bass_radix_sort", "bass_lift_counted",
                 "bass_set_count", "bass_counted_read", "bass_diag2q_counted",
                 "bass_merge_sorted", "bass_ht_insert", "bass_ht_lookup", "bass_workspace_bytes"]:
        assert must in names


def test_every_header_symbol_is_exported(lib):
    for name in header_functions():
        assert hasattr(lib, name), f"{name} not exported by libbass_hip.so"


def test_ctypes_table_matches_header():
    L = pkg_module("_lib")
    assert sorted(L.SIGNATURES) == header_functions()


def test_workspace_sizes(lib):
    a = lib.bass_workspace_bytes(1)
    b = lib.bass_workspace_bytes(1 << 20)
    c = lib.bass_workspace_bytes(1 << 22)
    assert 0 < a < b < c
    assert c - b >= 48 * ((1 << 22) - (1 << 20))


def test_argument_errors_without_gpu(lib):
    L = pkg_module("_lib")
    ht = ctypes.c_int64(0)
    G = (ctypes.c_double * 8)()
    rc = lib.bass_gate(None, None, 10, 3, 0, 1, G, 0, 1, 0, None, None, 10, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_ARG and "nq" in L.last_error()
    rc = lib.bass_gate(None, None, 10, 2, 4, 4, G, 0, 1, 0, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_ARG and "qubits" in L.last_error()
    rc = lib.bass_gate(None, None, 10, 1, 0, 0, G, 1, 0, 0, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_ARG and "rotation" in L.last_error()
    rc = lib.bass_gate(None, None, 10, 1, 0, 0, G, 0, 1, 1, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_OVERFLOW and "segmented" in L.last_error()
    rc = lib.bass_truncate(None, None, 5, 5, None, None, 5, 1, 0, 0, None, 0, None, 0, None)
    assert rc == L.BASS_E_ARG
    rc = lib.bass_truncate(None, None, 10, 5, None, None, 4, 1, 0, 0, None, 0, None, 0, None)
    assert rc == L.BASS_E_OVERFLOW
    rc = lib.bass_truncate(None, None, 10, 5, None, None, 5, 1, 0, 1, None, 0, None, 0, None)
    assert rc == L.BASS_E_ARG
    bad = (ctypes.c_int32 * 3)(0, 0, 2)
    rc = lib.bass_truncate(None, None, 10, 5, None, None, 5, 1, 0, 0, bad, 3, None, 0, None)
    assert rc == L.BASS_E_ARG and "permutation" in L.last_error()
    rc = lib.bass_lift(None, None, 10, 2, 0, G, 0, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_ARG and "bit" in L.last_error()
    rc = lib.bass_lift(None, None, 10, 2, 3, G, 2, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_ARG and "move" in L.last_error()
    rc = lib.bass_lift(None, None, 3000, 1, 3, G, 0, None, None, 2048, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_OVERFLOW
    # the no-scan flag (mode bit 4) leaves the mode check intact
    rc = lib.bass_lift(None, None, 10, 1, 3, G, 3 | L.LIFT_NO_SCAN, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_ARG and "mode" in L.last_error()
    rc = lib.bass_lift(None, None, 3000, 1, 3, G, L.LIFT_NO_SCAN, None, None, 2048, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_OVERFLOW
    rc = lib.bass_radix_sort(None, None, 10, 0, None, 0, None)
    assert rc == L.BASS_E_ARG
    rc = lib.bass_lift_counted(None, None, 10, 3, 3, G, 0, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_ARG and "nq" in L.last_error()
    rc = lib.bass_diag2q_counted(None, None, 10, 2, 2, G, None, 0, None)
    assert rc == L.BASS_E_ARG and "qubits" in L.last_error()
    rc = lib.bass_set_count(None, 0, -1, None)
    assert rc == L.BASS_E_ARG and "bad n" in L.last_error()
    rc = lib.bass_radix_sort(None, None, 1 << 30, 40, None, 0, None)
    assert rc == L.BASS_E_ARG and "2^30" in L.last_error()
    rc = lib.bass_sort_digit_bits(9)
    assert rc == L.BASS_E_ARG and "0, 8 or 10" in L.last_error()
    assert lib.bass_sort_digit_bits(10) == 0 and lib.bass_sort_digit_bits(0) == 10
    # digit plan: 8-bit digits unless 10 were requested
    P = ctypes.c_int(0)
    for bits, want in [(40, (8, 5)), (56, (8, 7)), (64, (8, 8)), (5, (8, 1))]:
        rb = lib.bass_sort_plan(bits, ctypes.byref(P))
        assert (rb, P.value) == want, bits
    lib.bass_sort_digit_bits(10)
    try:
        for bits, want in [(40, (10, 4)), (56, (10, 6)), (64, (10, 7))]:
            rb = lib.bass_sort_plan(bits, ctypes.byref(P))
            assert (rb, P.value) == want, bits
    finally:
        lib.bass_sort_digit_bits(0)
    assert lib.bass_sort_plan(0, ctypes.byref(P)) == L.BASS_E_ARG
    rc = lib.bass_dense_gate(None, 0, 1, 0, 0, G, None)
    assert rc == L.BASS_E_ARG and "N out of range" in L.last_error()
    rc = lib.bass_dense_gate(None, 4, 3, 0, 1, G, None)
    assert rc == L.BASS_E_ARG and "nq" in L.last_error()
    rc = lib.bass_dense_gate(None, 4, 2, 2, 2, G, None)
    assert rc == L.BASS_E_ARG and "qubits" in L.last_error()
    bad3 = (ctypes.c_int32 * 3)(0, 0, 1)
    rc = lib.bass_dense_scatter(None, None, 4, bad3, 3, None, 3, None)
    assert rc == L.BASS_E_ARG and "permutation" in L.last_error()
    rc = lib.bass_dense_scatter(None, None, 9, None, 0, None, 3, None)
    assert rc == L.BASS_E_ARG and "bad n" in L.last_error()
    rho = (ctypes.c_double * 32)()
    rc = lib.bass_rdm2q_ordered(None, None, 10, 0, rho, None, 0, None)
    assert rc == L.BASS_E_ARG and "bit j" in L.last_error()
    rc = lib.bass_rdm2q_ordered(None, None, 10, 3, rho, None, 0, None)
    assert rc == L.BASS_E_WORKSPACE
    rc = lib.bass_gate(None, None, 10, 1, 0, 0, G, 0, 1, 0, None, None, 40, ctypes.byref(ht), None, 0, None)
    assert rc == L.BASS_E_WORKSPACE


def test_errors_map_to_reference_exceptions():
    L = pkg_module("_lib")
    with pytest.raises(RuntimeError, match="Temporary buffer overflow"):
        L.check(L.BASS_E_OVERFLOW, "x")
    with pytest.raises(L.BassError):
        L.check(L.BASS_E_DEVICE, "x")


def test_no_gpu_means_loud_failure():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    dev = pkg_module("_device")
    with pytest.raises(pkg_module("_lib").BassError):
        dev.context("cuda:0")


def test_job_table_layout():
    """bass_job as the ctypes table sees it: 616 bytes, the count cells and the
    per-job gate width last."""
    L = pkg_module("_lib")
    assert L.JOB_DTYPE.itemsize == 616
    assert L.JOB_DTYPE.fields["cnt_in"][1] == 592 and L.JOB_DTYPE.fields["cnt_out"][1] == 600
    assert L.JOB_DTYPE.fields["nq"][1] == 608


def test_batch_plan_arguments_without_gpu(lib):
    L = pkg_module("_lib")
    a = lib.bass_bplan_ws_bytes(8, 16)
    b = lib.bass_bplan_ws_bytes(8, 32)
    c = lib.bass_bplan_ws_bytes(16, 32)
    assert 0 < a < b < c
    assert b - a >= 16 * 2048 * 32  # the select's candidate arrays per tile slot
    assert lib.bass_bplan_ws_bytes(0, 16) == 0
    h = ctypes.c_void_p()
    rc = lib.bass_bplan_create(0, 16, None, None, None, 0, 0, ctypes.byref(h))
    assert rc == L.BASS_E_ARG and "capacities" in L.last_error()
    rc = lib.bass_bplan_create(8, 16, None, None, None, 0, 0, ctypes.byref(h))
    assert rc == L.BASS_E_ARG and "null" in L.last_error()
    rc = lib.bass_bplan_run(None, 0, None, 1, None, None)
    assert rc == L.BASS_E_ARG and "plan" in L.last_error()
    assert lib.bass_bplan_destroy(None) == L.BASS_OK


def test_batch_plan_table_checks_without_gpu(lib):
    """bass_bplan_run validates its job table on the host before any device
    work: tile capacity, gate widths of a mixed step."""
    L = pkg_module("_lib")
    O = pkg_module("_ops")
    fake = ctypes.c_void_p(1 << 20)  # never dereferenced: the calls fail on the host
    h = ctypes.c_void_p()
    nb = lib.bass_bplan_ws_bytes(4, 2)
    rc = lib.bass_bplan_create(4, 2, fake, fake, fake, nb, 0, ctypes.byref(h))
    assert rc == L.BASS_OK and h.value
    try:
        J = O.new_jobs(2)
        J["x"] = J["a"] = J["xo"] = J["ao"] = 1 << 20
        J["k"] = J["cap_out"] = 4096
        J["n"] = 4096  # 4 one-qubit tiles per job: more than the plan's 2
        rc = lib.bass_bplan_run(h, L.BOP_GATE1, J.ctypes.data_as(ctypes.c_void_p), 2, None, None)
        assert rc == L.BASS_E_OVERFLOW and "tiles" in L.last_error()
        J["n"] = 1
        J["nq"] = [1, 3]
        rc = lib.bass_bplan_run(h, L.BOP_GATE12, J.ctypes.data_as(ctypes.c_void_p), 2, None, None)
        assert rc == L.BASS_E_ARG and "gate qubits" in L.last_error()
        rc = lib.bass_bplan_run(h, L.BOP_GATE1, J.ctypes.data_as(ctypes.c_void_p), 5, None, None)
        assert rc == L.BASS_E_ARG and "plan holds" in L.last_error()
    finally:
        lib.bass_bplan_destroy(h)
