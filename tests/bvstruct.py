"""Member lists of a C struct definition, for comparing the BV plugin contract
(struct _BVOps / struct _p_BV) between two headers: comments dropped, each
declarator expanded to (name, type) in order, whitespace and parameter names
normalized away, PETSCHEADER(...) kept as one member."""
import re


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _norm(s):
    s = re.sub(r"\s+", " ", s).strip()
    s = re.sub(r"\s*([*,()\[\]])\s*", r"\1", s)
    return s


def struct_members(text, tag):
    text = _strip_comments(text)
    m = re.search(r"struct\s+" + re.escape(tag) + r"\s*\{", text)
    if not m:
        raise ValueError(f"struct {tag} not found")
    i, depth = m.end(), 1
    j = i
    while depth:
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
        j += 1
    body = text[i:j - 1]
    out = []
    for decl in body.split(";"):
        decl = _norm(decl)
        if not decl:
            continue
        fp = re.match(r"(.*)\(\*(\w+)\)\((.*)\)$", decl)
        if fp:
            out.append([fp.group(2), "fn:" + fp.group(1) + "(" + fp.group(3) + ")"])
            continue
        hd = re.match(r"PETSCHEADER\((.*)\)$", decl)
        if hd:
            out.append(["PETSCHEADER", hd.group(1)])
            continue
        tm = re.match(r"((?:const )?(?:struct )?\w+)\s*(.*)$", decl)
        base, rest = tm.group(1), tm.group(2)
        for d in rest.split(","):
            d = d.strip()
            stars = len(d) - len(d.lstrip("*"))
            name = d.lstrip("*")
            arr = re.match(r"(\w+)(\[.*\])$", name)
            suffix = ""
            if arr:
                name, suffix = arr.group(1), arr.group(2)
            out.append([name, base + "*" * stars + suffix])
    return out
