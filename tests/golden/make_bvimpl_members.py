"""Writes bvimpl_members.json: the member lists of struct _BVOps and struct _p_BV
in SLEPc's include/slepc/private/bvimpl.h (the BV plugin contract), read as
text from the reference tree.  Run where the reference tree exists:

    python tests/golden/make_bvimpl_members.py /root/reference

The fixture is data (names and declared types in order); the parser that
produces it is the one tests/test_bv_plugin.py applies to this repository's
slepc_amd/host/slepcimpl.h."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from bvstruct import struct_members  # noqa: E402


def main(ref):
    text = open(os.path.join(ref, "include", "slepc", "private", "bvimpl.h")).read()
    out = {"source": "include/slepc/private/bvimpl.h (SLEPc 3.22)",
           "_BVOps": struct_members(text, "_BVOps"), "_p_BV": struct_members(text, "_p_BV")}
    with open(os.path.join(HERE, "bvimpl_members.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(len(out["_BVOps"]), "ops,", len(out["_p_BV"]), "members")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
