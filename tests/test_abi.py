"""The C ABI: libwzfh.so loads, exports every entry point include/wzfh.h declares, and the
host-only setup works without a device (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

import bs45_quantum_explorer_amd as wz

HEADER = Path(__file__).resolve().parents[1] / 'include' / 'wzfh.h'


def declared():
    text = re.sub(r'/\*.*?\*/', '', HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r'\b(wzfh_[a-z_]+)\s*\(', text)))


def test_header_declares_expected_entry_points():
    names = declared()
    for must in ('wzfh_init', 'wzfh_stream_cells', 'wzfh_complete', 'wzfh_search_cells', 'wzfh_npaf_check'):
        assert must in names
    assert set(names) == set(wz.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(wz.LIBPATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_config_default_matches_reference_defaults():
    c = wz.config(44, (1, 7, 8, 8))
    assert (c.m6, c.thm211b, c.thm212, c.outer_k, c.ab_prof, c.abp_cap) == (1, 1, 1, 3, -1, 20000)
    c = wz.config(19, (0, 2, 5, 7))
    assert (c.m6, c.thm211b, c.thm212) == (0, 0, 0)


def test_bad_signature_is_rejected():
    with pytest.raises(wz.WzfhError) as e:
        wz.Problem(10, (1, 1, 1, 1), device=False)
    assert e.value.code == wz.EINVAL


def test_host_only_handle_refuses_device_calls():
    P = wz.Problem(6, (5, 1, 0, 0), device=False, m6=1)
    assert P.num_cells() > 0
    with pytest.raises(wz.WzfhError) as e:
        P.stream([0])
    assert e.value.code == wz.ENODEV
    with pytest.raises(wz.WzfhError) as e:
        P.tuning()
    assert e.value.code == wz.ENODEV


def test_host_npaf_and_flat_score_helpers():
    # BS(7,6) found by the reference's own first run (wz_match 6 5 1 0 0)
    A = [1, 1, 1, 1, 1, -1, 1]; B = [1, -1, 1, -1, 1, 1, -1]
    C = [1, 1, 1, -1, -1, -1]; D = [1, -1, -1, 1, 1, -1]
    assert wz.npaf_max(6, A, B, C, D) == 0
    assert wz.flat_score(6, wz.pack(C), wz.pack(D)) == 10   # FIRSTHIT score=10 in that run
