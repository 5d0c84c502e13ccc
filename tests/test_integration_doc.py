"""The Rust bindings written out for maintainers match the C header: every `pub fn gc_*` in
INTEGRATION.md is declared in include/gcompress.h with the same number of parameters, and each
parameter is a pointer on both sides or on neither (a u64 where the header has a pointer is the
mistake this catches)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _split_params(s):
    s = " ".join(s.split())
    if s in ("", "void"):
        return []
    return [p.strip() for p in s.split(",") if p.strip()]


def header_params():
    txt = open(os.path.join(ROOT, "include", "gcompress.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(gc_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", txt):
        out[m.group(1)] = [("*" in p) for p in _split_params(m.group(2))]
    return out


def rust_params():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    out = {}
    for block in re.findall(r"```rust\n(.*?)```", txt, flags=re.S):
        block = re.sub(r"//[^\n]*", "", block)
        for m in re.finditer(r"pub fn (gc_[a-z0-9_]+)\s*\(([^)]*)\)", block):
            out[m.group(1)] = [("*mut" in p or "*const" in p) for p in _split_params(m.group(2))]
    return out


def test_rust_bindings_match_header():
    h, r = header_params(), rust_params()
    assert len(r) >= 10
    for name, ptrs in r.items():
        assert name in h, f"INTEGRATION.md binds {name}, which gcompress.h does not declare"
        assert len(ptrs) == len(h[name]), f"{name}: {len(ptrs)} parameters in INTEGRATION.md, {len(h[name])} in the header"
        assert ptrs == h[name], f"{name}: pointer parameters differ (doc {ptrs}, header {h[name]})"


def test_checker_catches_a_wrong_binding():
    # the 6-parameter binding once written for gc_batch_copy_zstd fails the comparison
    h = header_params()
    wrong = "b: *mut GcBatch, i: u32, seqs: *mut GcZstdSequence, cap: u64, counts: *mut u32, total: *mut u64"
    ptrs = [("*mut" in p or "*const" in p) for p in _split_params(wrong)]
    assert ptrs != h["gc_batch_copy_zstd"]
