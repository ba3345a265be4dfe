"""Every code block of INTEGRATION.md matches the ABI: C and C++ blocks
compile against include/pagani.h with warnings as errors, and the Python
block calls Context methods with arguments their signatures accept."""
import ast
import inspect
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")


def blocks(lang):
    text = open(DOC).read()
    return re.findall(r"```" + lang + r"\n(.*?)```", text, flags=re.S)


@pytest.mark.parametrize("lang,cc,std", [("c", "gcc", "-std=c11"), ("cpp", "g++", "-std=c++17")])
def test_c_blocks_compile(tmp_path, lang, cc, std):
    found = blocks(lang)
    assert found, f"no {lang} blocks in INTEGRATION.md"
    for k, src in enumerate(found):
        f = tmp_path / f"snippet{k}.{lang}"
        f.write_text(src)
        r = subprocess.run([cc, std, "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                            "-I", os.path.join(ROOT, "include"), str(f)], capture_output=True, text=True)
        assert r.returncode == 0, f"{lang} block {k} does not compile:\n{src}\n{r.stderr}"


def test_python_block_matches_binding():
    import pagani_amd
    found = blocks("python")
    assert found
    for src in found:
        tree = ast.parse(src)
        calls = [n for n in ast.walk(tree) if isinstance(n, ast.Call) and isinstance(n.func, ast.Attribute)
                 and isinstance(n.func.value, ast.Name) and n.func.value.id == "ctx"]
        assert calls
        for c in calls:
            meth = getattr(pagani_amd.Context, c.func.attr, None)
            assert meth is not None, f"Context has no method {c.func.attr}"
            sig = inspect.signature(meth)
            sig.bind(None, *[None] * len(c.args), **{kw.arg: None for kw in c.keywords})
        for n in ast.walk(tree):
            if isinstance(n, ast.Attribute) and isinstance(n.value, ast.Name) and n.value.id == "pagani_amd":
                assert hasattr(pagani_amd, n.attr), f"pagani_amd has no {n.attr}"
