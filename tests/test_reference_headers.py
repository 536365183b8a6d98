"""The five public type headers against the reference's own headers, compiled.

For every struct of include/*_types.h a small C program prints the struct's
size and each field's offset and size, and every enumerator's value; it is
compiled with gcc once against this repository's header and once against the
reference's header (with an empty PSLP_stats.h, the only include the
reference's LP header needs that is not in the tree), and the two outputs
must be identical.  Runs where the reference tree is present (a CPU test; it
reads the reference headers as text through the compiler and runs nothing of
the reference's)."""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")

PAIRS = [  # (our header, reference include directory, reference header name)
    ("cp_al_types.h", "LP/CP_AL/include", "cp_al_types.h"),
    ("fa_cp_types.h", "QP/FA_CP/include", "fa_cp_types.h"),
    ("lin_cp_al_types.h", "QP/LIN_CP_AL/include", "cp_al_types.h"),
    ("planted_cp_al_types.h", "QP/PLANTED_CP_AL/include", "cp_al_types.h"),
    ("planted_fa_cp_types.h", "QP/PLANTED_FA_CP/include", "fa_cp_types.h"),
]


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _match_brace(text: str, i: int) -> int:
    """Index of the '}' closing the '{' at text[i]."""
    depth = 0
    for j in range(i, len(text)):
        depth += text[j] == "{"
        depth -= text[j] == "}"
        if depth == 0:
            return j
    raise ValueError("unbalanced braces")


def _fields(body: str, prefix: str = ""):
    """Member paths of a struct body, nested unions / structs as dotted paths."""
    out, i = [], 0
    while i < len(body):
        m = re.match(r"\s*(?:union|struct)\s*\w*\s*\{", body[i:])
        if m:
            lb = i + m.end() - 1
            rb = _match_brace(body, lb)
            semi = body.index(";", rb)
            name = body[rb + 1:semi].strip()
            out.append(prefix + name)
            out += _fields(body[lb + 1:rb], prefix + name + ".")
            i = semi + 1
            continue
        semi = body.find(";", i)
        if semi < 0:
            break
        decl = body[i:semi].strip()
        i = semi + 1
        if not decl or "(" in decl:
            continue
        head, *rest = [p.strip() for p in decl.split(",")]  # "type a, b" / "type *a" / "type a[4]"
        names = [re.sub(r"\[.*$", "", head.split()[-1]).lstrip("*")]
        names += [re.sub(r"\[.*$", "", r).lstrip("*") for r in rest]
        out += [prefix + n for n in names]
    return out


def _structs(text: str):
    out = []
    for m in re.finditer(r"typedef\s+struct\s*\w*\s*\{", text):
        lb = m.end() - 1
        rb = _match_brace(text, lb)
        name = re.match(r"\s*(\w+)\s*;", text[rb + 1:]).group(1)
        out.append((name, _fields(text[lb + 1:rb])))
    return out


def _enums(text: str):
    out = []
    for body in re.findall(r"typedef\s+enum\s*\w*\s*\{(.*?)\}\s*\w+\s*;", text, flags=re.S):
        for item in body.split(","):
            item = item.strip()
            if item:
                out.append(item.split("=")[0].strip())
    return out


def _program(include: str, structs, enums) -> str:
    lines = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{include}"', "int main(void) {"]
    for name, fields in structs:
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'  printf("{name}.{f} %zu %zu\\n", offsetof({name}, {f}), sizeof((({name} *)0)->{f}));')
    for e in enums:
        lines.append(f'  printf("{e} %ld\\n", (long){e});')
    lines.append("  return 0;\n}")
    return "\n".join(lines) + "\n"


def _run(tmp: Path, tag: str, src: str, incdirs) -> str:
    c = tmp / f"{tag}.c"
    exe = tmp / tag
    c.write_text(src)
    subprocess.run(["gcc", "-std=c11", *[f"-I{d}" for d in incdirs], str(c), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


@pytest.mark.skipif(not REF.exists() or shutil.which("gcc") is None, reason="reference tree or gcc absent")
@pytest.mark.parametrize("ours, refdir, refname", PAIRS)
def test_layout_matches_compiled_reference_header(tmp_path, ours, refdir, refname):
    text = _strip_comments((ROOT / "include" / ours).read_text())
    structs, enums = _structs(text), _enums(text)
    assert structs and enums
    (tmp_path / "stub").mkdir()
    (tmp_path / "stub" / "PSLP_stats.h").write_text("/* empty: no public struct uses a PSLP type */\n")
    out_ours = _run(tmp_path, "ours_prog", _program(ours, structs, enums), [ROOT / "include"])
    out_ref = _run(tmp_path, "ref_prog", _program(refname, structs, enums), [REF / refdir, tmp_path / "stub"])
    assert out_ours == out_ref
    assert out_ours.count("\n") >= sum(1 + len(f) for _, f in structs) + len(enums)
