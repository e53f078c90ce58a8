"""libtp.so loads on a machine without a GPU and exports every entry point
that include/*.h declares; GPU entry points fail cleanly when no GPU exists."""
import ctypes as C
import os
import re

import pytest

from conftest import PKG, ROOT


def declared_functions():
    names = set()
    for h in ("tp.h", "tp_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**\s*(tp_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(PKG, "build", "libtp.so"))
    names = declared_functions()
    assert len(names) >= 20, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import trois_photons as tp
    with pytest.raises(tp.TpError, match="no usable GPU"):
        tp.Engine()


def test_null_and_bad_arguments():
    import trois_photons as tp
    lib = tp.lib
    assert lib.tp_create(0, None) == -1
    assert lib.tp_destroy(None) == -1
    assert lib.tp_rng_init(7, C.byref(tp.XoshiroState())) == -1
    assert lib.tp_rng_skip(tp.RNG_RANF, None, 3) == -1
    st = tp.RanfState()
    st.index = 99
    assert lib.tp_rng_skip(tp.RNG_RANF, C.byref(st), 3) == -1
