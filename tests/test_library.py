"""CPU checks of the C ABI library: it loads without a GPU and exports every
entry point that include/libp_amd.h declares."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "libp_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(libp_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "libp_ax_hex_partial_f64" in syms and "libp_ogs_gather_f64" in syms
    assert len(syms) > 40


def test_library_exports_every_declared_symbol():
    import libparanumal_amd as lp
    L = lp.lib()  # loads libp_amd.so and binds every signature
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(lp.SIGNATURES), set(declared_symbols()) ^ set(lp.SIGNATURES)


def test_status_strings_and_argument_errors_without_gpu():
    import libparanumal_amd as lp
    L = lp.lib()
    assert L.libp_status_string(0) == b"ok"
    # degree out of range is rejected before any device work
    assert L.libp_ax_hex_partial_f64(0, 1, None, None, None, None, None, 0.0, None, None, None) != 0
    assert b"degree" in L.libp_last_error() or b"null" in L.libp_last_error()
    assert L.libp_mg_coarsen_hex_f64(3, 3, 1, None, None, None, None, None, None) != 0
