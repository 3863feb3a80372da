"""The Rust binding source (crates/eplace-sys, not compiled in this image:
no Rust toolchain) against the C header: every #[repr(C)] struct has the
header's fields in the header's order with the same sizes, every function
the extern block declares is one the header declares and the library
exports, and the glue calls only declared functions."""
import re
from pathlib import Path

import ctypes as C

import eisenjoch_amd._lib as L

ROOT = Path(__file__).resolve().parent.parent
RS = (ROOT / "crates" / "eplace-sys" / "src" / "lib.rs").read_text()
GLUE = (ROOT / "crates" / "eplace-sys" / "glue" / "electro_place_gpu.rs").read_text()

SIZES = {"i32": 4, "u32": 4, "i64": 8, "u64": 8, "f64": 8, "f32": 4, "u8": 1}
CT = {"ep_design": L.EpDesign, "ep_config": L.EpConfig, "ep_result": L.EpResult,
      "ep_iter_rec": L.EpIterRec, "ep_legal_problem": L.EpLegalProblem,
      "ep_legal_result": L.EpLegalResult}


def rust_structs():
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+) \{(.*?)\n\}", RS, re.S):
        fields = re.findall(r"pub (\w+): ([^,\n]+),", m.group(2))
        out[m.group(1)] = fields
    return out


def test_rust_structs_mirror_the_header():
    rs = rust_structs()
    for name, py in CT.items():
        assert name in rs, name
        rf = rs[name]
        assert len(rf) == len(py._fields_), name
        for (rn, rt), (pn, pt) in zip(rf, py._fields_):
            assert rn.lower() == pn.rstrip("_").lower(), (name, rn, pn)
            size = 8 if rt.strip().startswith("*") else SIZES[rt.strip()]
            assert size == C.sizeof(pt), (name, rn, rt)


def test_rust_extern_block_is_the_header():
    decl = set(re.findall(r"pub fn (ep_\w+)\(", RS))
    assert decl <= set(L.SIGNATURES), decl - set(L.SIGNATURES)
    lib = L.lib()
    for f in decl:
        assert hasattr(lib, f), f
    # the glue reaches the library only through the safe layer or declared functions
    used = set(re.findall(r"ffi::(ep_\w+)\(", GLUE))
    assert used <= decl
    assert "place_legalized" in GLUE and "flatten_for_gpu" in GLUE
