"""Writes tests/golden/abi_layout.json: the boundary layout of the
reference's own header (/root/reference/c_lib/nr_ldpc.h) as compiled by gcc
on this host (tests/abi/abi_probe.c).  Run in the build container only;
tests/test_abi_layout.py compiles the same probe against
include/nr_ldpc_mi355.h and compares."""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/c_lib"


def probe(include_dir, header):
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "probe")
        subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", f"-I{include_dir}",
                        f'-DABI_HEADER="{header}"', "-o", exe,
                        os.path.join(ROOT, "tests", "abi", "abi_probe.c"),
                        "-Wl,--unresolved-symbols=ignore-all"], check=True)
        return json.loads(subprocess.run([exe], check=True, capture_output=True,
                                         text=True).stdout)


if __name__ == "__main__":
    layout = probe(REF, "nr_ldpc.h")
    layout["source"] = "c_lib/nr_ldpc.h:1-131 compiled by gcc (x86-64 LP64) with tests/abi/abi_probe.c"
    with open(os.path.join(HERE, "abi_layout.json"), "w") as f:
        json.dump(layout, f, indent=1, sort_keys=True)
    print(json.dumps(layout, indent=1, sort_keys=True))
