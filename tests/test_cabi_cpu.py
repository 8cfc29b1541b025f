"""The C ABI library loads and exports every function include/cc_hip.h
declares, and the binding names them all (no GPU call is made)."""
import os
import re
import sys

import golden_util as gu

sys.path.insert(0, os.path.join(gu.REPO, "cosmic-collisions_amd"))
import ccamd  # noqa: E402


def declared():
    text = open(os.path.join(gu.REPO, "include", "cc_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(cc_\w+)\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared()
    assert "cc_step" in names and "cc_slots" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = ccamd.lib()
    for name in declared():
        assert hasattr(L, name), name


def test_binding_covers_every_symbol():
    assert sorted(ccamd.EXPORTS) == declared()
