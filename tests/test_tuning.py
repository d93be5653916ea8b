"""The per-size-class tuning table (CPU): schema of the committed table,
size classes, and rejection of values the library does not allow."""

import pytest

from conftest import pkg_module


def test_size_classes():
    T = pkg_module("tuning")
    assert [T.size_class(n) for n in (1, (1 << 14) - 1, 1 << 14, (1 << 18) - 1, 1 << 18, 1 << 22, 1 << 30)] == \
        [0, 0, 1, 1, 2, 3, 3]


def test_committed_table_is_valid():
    T = pkg_module("tuning")
    table = T.load_table()
    if table is None:
        pytest.skip("no tuning table committed")
    T.validate(table)
    assert len(table["representative_n"]) == 4
    for name in T.KNOBS:
        assert all(v in T.VALUES[name] for v in table[name]), name


def test_validate_rejects_bad_values():
    T = pkg_module("tuning")
    good = {k: [0, 0, 0, 0] for k in T.KNOBS}
    T.validate(good)
    for name, bad in (("table_factor", 3), ("emit_items_1q", 1), ("emit_items_2q", 8), ("sort_digits", 9)):
        t = {k: list(v) for k, v in good.items()}
        t[name][2] = bad
        with pytest.raises(ValueError):
            T.validate(t)
    with pytest.raises(ValueError):
        T.validate({**good, "sort_digits": [8, 8]})
