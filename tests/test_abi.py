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
