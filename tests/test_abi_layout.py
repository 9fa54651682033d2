"""The boundary structs and tables of include/nr_ldpc_mi355.h, compiled by a
C compiler, against the reference's c_lib/nr_ldpc.h:42-131 compiled the same
way (tests/golden/abi_layout.json, written by tests/golden/gen_abi_layout.py):
sizeof, offsetof and field sizes of ldpc_in_config_t, ldpc_config_t and
ldpc_decoder_cfg_t, the sizes of the exported base-graph tables and
HARQ_memory, the header constants, and -- by compiling function-pointer
assignments with -Werror -- the types of the six functions the reference's
Python callers bind."""
import json
import os
import shutil
import sys

import pytest

from conftest import GOLDEN, ROOT

sys.path.insert(0, GOLDEN)
from gen_abi_layout import probe  # noqa: E402


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_header_layout_equals_reference():
    with open(os.path.join(GOLDEN, "abi_layout.json")) as f:
        want = json.load(f)
    want.pop("source")
    got = probe(os.path.join(ROOT, "include"), "nr_ldpc_mi355.h")
    assert got == want
