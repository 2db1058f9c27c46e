"""Regenerate tests/golden/case_headers.json: the values of the reference's
FP64 case headers as a C++ compiler evaluates them.

Each `__constant__ T name = expr;` line of a header (and its `#define PI`)
becomes a `constexpr T name = expr;` in a throw-away C++ file that g++
compiles and runs, printing every value as a hexadecimal float; the JSON
keeps the exact doubles.  The fixture is therefore independent of the
project's own header parser (hydist_amd/case.py), which the test in
tests/test_case_parser.py checks against it.

Run from the repository root, with the reference checkout at /root/reference:

    python tests/golden/make_case_headers.py
"""
import json
import os
import re
import subprocess
import sys
import tempfile

REF = "/root/reference"
HEADERS = ("constant.cuh", "constant_TanChau.cuh", "constant_songluy.cuh", "constant_songluy_980x1048.cuh")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "case_headers.json")


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def program(text):
    text = strip_comments(text)
    pi = re.search(r"#define\s+PI\s+([0-9.eE+-]+)", text).group(1)
    decls = re.findall(r"__constant__\s+(int|DOUBLE|double)\s+(\w+)\s*=\s*([^;]+);", text)
    lines = ["#include <cstdio>", f"#define PI {pi}", "#define DOUBLE double"]
    body = [f'    std::printf("PI %a\\n", (double)(PI));']
    for ty, name, expr in decls:
        cty = "int" if ty == "int" else "double"
        lines.append(f"constexpr {cty} v_{name} = {expr};")
        fmt = "%d" if cty == "int" else "%a"
        body.append(f'    std::printf("{name} {fmt}\\n", v_{name});')
    return "\n".join(lines + ["int main()", "{"] + body + ["    return 0;", "}", ""])


def evaluate(path):
    with open(path) as f:
        src = program(f.read())
    with tempfile.TemporaryDirectory() as d:
        cpp, exe = os.path.join(d, "h.cpp"), os.path.join(d, "h")
        with open(cpp, "w") as f:
            f.write(src)
        subprocess.run(["g++", "-std=c++17", "-O0", "-o", exe, cpp], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    vals = {}
    for line in out.splitlines():
        name, v = line.split()
        vals[name] = int(v) if name == "segment_limit" else float.fromhex(v)
    return dict(sorted(vals.items()))


def main():
    if not os.path.isdir(REF):
        sys.exit(f"{REF} not present")
    golden = {h: evaluate(os.path.join(REF, h)) for h in HEADERS}
    with open(OUT, "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {OUT}: " + ", ".join(f"{h} ({len(v)} values)" for h, v in golden.items()))


if __name__ == "__main__":
    main()
