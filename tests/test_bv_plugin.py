"""The BV plugin contract (CPU): BVHIPKS is written against struct _BVOps and
struct _p_BV with exactly the members, order and declared types of the
reference's include/slepc/private/bvimpl.h:25-113 (the fixture
tests/golden/bvimpl_members.json, read from that header by
tests/golden/make_bvimpl_members.py), so the type compiles into an unmodified
SLEPc; it fills only those ops, including the gramschmidt hook that the
reference's BVOrthogonalizeGS dispatches to (bvorthog.c:134); everything else
it offers goes through the one composed function "BVGetKrylovOps_C"."""
import json
import os
import re
import subprocess

import pytest

from bvstruct import struct_members

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "slepc_amd", "host")


@pytest.fixture(scope="module")
def ref():
    return json.load(open(os.path.join(ROOT, "tests", "golden", "bvimpl_members.json")))


@pytest.fixture(scope="module")
def impl_text():
    return open(os.path.join(HOST, "slepcimpl.h")).read()


def test_bvops_members_match_reference(ref, impl_text):
    assert struct_members(impl_text, "_BVOps") == ref["_BVOps"]
    assert len(ref["_BVOps"]) == 35 and ["gramschmidt", "fn:PetscErrorCode(BV,PetscInt,Vec,PetscBool*,PetscScalar*,"
                                         "PetscScalar*,PetscReal*,PetscReal*)"] in ref["_BVOps"]


def test_p_bv_members_match_reference(ref, impl_text):
    ours = struct_members(impl_text, "_p_BV")
    assert ours == ref["_p_BV"]
    assert ours[0] == ["PETSCHEADER", "struct _BVOps"] and ours[-1] == ["data", "void*"]


def test_parser_sees_a_changed_member(impl_text):
    # a renamed, an added and a retyped member all change the list
    base = struct_members(impl_text, "_p_BV")
    assert struct_members(impl_text.replace("PetscInt           nc;", "PetscInt           ncx;"), "_p_BV") != base
    assert struct_members(impl_text.replace("void               *data;", "void *data; int extra;"), "_p_BV") != base
    assert struct_members(impl_text.replace("PetscReal          deftol;", "double deftol;"), "_p_BV") != base


def _ops_filled(src, fn):
    body = src[src.index(fn):]
    body = body[:body.index("\n}\n")]
    return re.findall(r"bv->ops->(\w+)\s*=", body)


def test_hipks_fills_reference_ops_only(ref):
    src = open(os.path.join(HOST, "bvhipks.c")).read()
    filled = _ops_filled(src, "PetscErrorCode BVCreate_HIPKS(BV bv)")
    names = {n for n, _ in ref["_BVOps"]}
    assert filled and set(filled) <= names
    assert "gramschmidt" in filled and "dot" in filled and "multinplace" in filled
    # the only other channel is the one composed function
    assert re.findall(r'PetscObjectComposeFunction\(bv, "(\w+)"', src) == ["BVGetKrylovOps_C", "BVGetKrylovOps_C"]


def test_plugin_compiles_against_the_header():
    # the type's file alone, warnings as errors: every bv-> access is a member of
    # the reference layout, since the header's structs have no other members
    for f in ("bvhipks.c", "mat.c"):
        r = subprocess.run(["gcc", "-fsyntax-only", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-Wall", "-Werror",
                            "-Wno-unused-function", os.path.join(HOST, f)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
