"""The C-ABI library loads on the CPU and exports every symbol the header declares."""
import os
import re

import gcompress

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "gcompress.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(gc_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_header():
    L = gcompress.lib()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), f"{n} declared in gcompress.h but not exported"


def test_python_signatures_cover_header():
    declared = set(header_functions())
    bound = {n for n, _, _ in gcompress.SIGNATURES}
    assert declared == bound, (declared - bound, bound - declared)


def test_last_error_is_a_string():
    assert isinstance(gcompress.lib().gc_last_error(), bytes)


def test_constants_match_header():
    txt = open(os.path.join(ROOT, "include", "gcompress.h")).read()
    m = re.search(r"#define GC_SEG_DEFAULT_MAX_SUB (\d+)u", txt)
    assert m and int(m.group(1)) == gcompress.SEG_DEFAULT_MAX_SUB


def test_container_output_buffer_checked():
    """An output buffer below the container bound is refused before any device work."""
    import numpy as np
    import pytest
    x = np.zeros(1000, np.uint8)
    with pytest.raises(ValueError):
        gcompress.nvzs_compress_gpu(x, 1, out=np.zeros(10, np.uint8))
    with pytest.raises(ValueError):
        gcompress.nvlz_compress_gpu(x, 128, out=np.zeros(gcompress.lib().gc_container_bound(1000, 1), np.uint32))
