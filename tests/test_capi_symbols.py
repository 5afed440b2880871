"""The C-ABI library loads and exports every function the public headers
declare.  No compute calls (no GPU needed)."""
import re
from pathlib import Path

import ctypes

from tgtp_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(tgtp_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("tgtp_score", "tgtp_saba_walks", "tgtp_taus", "tgtp_ctx_create",
                 "tgtp_graph_upload", "tgtp_load_txt"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTED) == sorted(declared_functions())


def test_strerror_without_device():
    lib = _lib.lib()
    assert lib.tgtp_strerror(_lib.TGTP_E_ARG) == b"invalid argument"
