"""Field constants written into the device headers, checked against their definitions."""
import os
import re

from helpers import O

CSRC = os.path.join(os.path.dirname(__file__), "..", "weighted_bte_amd", "csrc")


def _macro_limbs(text, name):
    m = re.search(r"#define %s\s*\\?\s*((?:0x[0-9a-f]+u,?\s*\\?\s*)+)" % name, text)
    assert m, name
    return [int(x, 16) for x in re.findall(r"0x([0-9a-f]+)u", m.group(1))]


def test_fp_header_constants():
    text = open(os.path.join(CSRC, "fp.hpp")).read()
    val = lambda limbs: sum(v << (32 * i) for i, v in enumerate(limbs))
    p = val(_macro_limbs(text, "BTE_P_LIMBS"))
    assert p == O.P
    assert val(_macro_limbs(text, "BTE_P2_LIMBS")) == O.P * O.P
    assert val(_macro_limbs(text, "BTE_ONE_LIMBS")) == (1 << 384) % O.P
    assert val(_macro_limbs(text, "BTE_R2_LIMBS")) == (1 << 768) % O.P
