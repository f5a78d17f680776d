"""The C ABI: every function include/jurassic_amd.h declares is exported by
libjurassic_amd.so and bound in pyjurassic.lib, with the same arity; the
struct layouts match the reference's.  Loading the library and looking up
symbols needs no GPU (nothing is called)."""
import re
from pathlib import Path

from pyjurassic import jtypes as J
from pyjurassic import lib as L

HEADER = Path(__file__).resolve().parents[1] / "include" / "jurassic_amd.h"


def declared():
    """(name, number of parameters) of every function prototype."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    text = re.sub(r"typedef struct \{.*?\} \w+;", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(1)] = len(args)
    return out


def test_header_parses():
    d = declared()
    assert "formod" in d and d["formod"] == 4
    assert d["formod_pencil"] == 6
    assert d["jur_formod_batch_device"] == 5
    assert len(d) >= 22


def test_every_declared_symbol_is_exported_and_bound():
    lib = L.load()
    for name, nargs in declared().items():
        assert hasattr(lib, name), name            # dlsym succeeds
        assert name in L.SIGNATURES, name
        assert len(L.SIGNATURES[name][1]) == nargs, name
    assert set(L.SIGNATURES) == set(declared())


def test_struct_layouts():
    J.check_layout()
    import ctypes as C
    # 3 ints, pad, 19 pointers, out_ld (+ pad)
    assert C.sizeof(L.jur_batch_t) == 3 * 4 + 4 + 19 * 8 + 8
    assert L.jur_batch_t.out_ld.offset == 3 * 4 + 4 + 19 * 8
    # gsl_matrix: three size_t, two pointers, int (+ padding)
    assert C.sizeof(L.jur_gsl_matrix) == 48 and L.jur_gsl_matrix.data.offset == 24
