"""The C ABI: libgle.so loads and exports every symbol include/gle.h declares."""
import os
import re

from garli_amd import engine as E

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gle.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(gle_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = E.lib()
    names = declared()
    assert len(names) > 40
    for n in names:
        assert hasattr(lib, n), n


def test_bindings_cover_header():
    assert sorted(E.SIGNATURES) == declared()


def _create(**kw):
    d = dict(n_tips=8, n_patterns=100, n_states=4, n_rate_cats=4, n_eigen=1, n_cla_slots=20, n_pmat_slots=20,
             precision=0)
    d.update(kw)
    h = E.C.c_int(-1)
    return E.lib().gle_create_instance(E.C.byref(E.InstanceDesc(**d)), E.C.byref(h))


def test_instance_description_checked_before_the_device():
    """Empty, ragged-impossible and oversized descriptions are refused with
    GLE_ERR_ARG (or GLE_ERR_UNSUPPORTED for a state count without kernels)
    before any device work."""
    for kw in (dict(n_patterns=0), dict(n_tips=1), dict(n_rate_cats=0), dict(n_rate_cats=9),
               dict(n_eigen=5), dict(n_cla_slots=0), dict(n_pmat_slots=0), dict(precision=7), dict(n_states=65)):
        assert _create(**kw) == -1, kw
    assert _create(n_states=7) == -5
    assert _create(n_states=3) == -5
    if E.device_count() == 0:  # no GPU here: a valid description fails on the device
        assert _create() == -2
        assert b"device" in E.lib().gle_last_error().lower() or b"gpu" in E.lib().gle_last_error().lower()


def test_level_ranges_checked_before_the_device():
    """gle_update_partials refuses level ranges that skip, reorder or overrun
    the op list, and levels whose ops read a CLA written beside them, before
    it looks at the instance (so this runs without a GPU)."""
    ops = (E.ClaOp * 3)(E.ClaOp(0, -1, 0, -2, 1, 0), E.ClaOp(1, -3, 2, -4, 3, 0), E.ClaOp(2, 0, 4, 1, 5, 1))
    lib = E.lib()

    def call(levels):
        ls = (E.C.c_int * len(levels))(*levels)
        return lib.gle_update_partials(12345, ops, 3, ls, len(levels) - 1), lib.gle_last_error().decode()

    for bad, what in (([0, 4, 3], "range"), ([0, 2, 1, 3], "range"), ([1, 3], "cover"), ([0, 3], "own level"),
                      ([0, -1, 3], "range")):
        rc, msg = call(bad)
        assert rc == -1 and what in msg, (bad, msg)
    rc, msg = call([0, 2, 3])  # valid levels: now the bad handle is what fails
    assert rc == -1 and "instance" in msg
