#!/usr/bin/env python3
"""Extract the known-answer vectors held in the reference's test sources into JSON fixtures.

Run in a container where the reference tree is mounted (default /root/reference).
It reads the C array literals as text and writes only data (inputs and expected
outputs) into tests/golden/*.json; no reference code is copied or executed.

Sources (reference = lCardosoSantos/FK20-CUDA):
  fr_fft_testvector.cu:10-1038  q[512] -> a[512] = fr_fft(q)       (used by frtest_fft.cu:35-86)
  fk20test_kat.cu:31,1317       g1p_input[512] -> g1p_correct[512] = g1p_fft(g1p_input)
                                (checked by g1p_fft_verify / g1p_ift_verify, fk20test_kat.cu:3371-3444)
  fr_roots.cu:17-533            fr_roots[515] twiddle table
  g1test_kat.cu:10-60           0, 2G, 3G, 24G, 25G, wG, w^8G (G1P_ANYINF = 1 variant of g1p_x0)
  fptest_kat.cu:18-716          per-operation Fp KATs (cpy, reduce6, eq/neq, neg, x2, x3, add,
                                sub, sqr, mul, inv, Mandelbrot)
  frtest_kat.cu:15-650          per-operation Fr KATs (cpy, reduce4, eq/neq, neg, x2, x3, add,
                                sub, sqr, mul, inv)
Each value is written as a big-endian hex string of the integer the limbs encode.
"""
import json
import os
import re
import sys

REF = os.environ.get("FK20_REFERENCE_DIR", "/root/reference")
OUT = os.path.dirname(os.path.abspath(__file__))

NUM = re.compile(r"0[xX][0-9a-fA-F]+|\b\d+\b")


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def keep_if_branches(text):
    """Keep the #if branch of #if/#else/#endif blocks (G1P_ANYINF is 1 in g1.cuh:14)."""
    out, skip, depth = [], False, 0
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("#if"):
            depth += 1
            continue
        if s.startswith("#else"):
            skip = True
            continue
        if s.startswith("#endif"):
            depth -= 1
            skip = False
            continue
        if not skip:
            out.append(line)
    return "\n".join(out)


def brace_block(text, start):
    """Return the text of the balanced {...} block starting at text[start] == '{'."""
    assert text[start] == "{"
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start : i + 1]
    raise ValueError("unbalanced braces")


def leaves(block):
    """Innermost {...} groups of a nested initializer, each as a list of ints."""
    groups = re.findall(r"\{([^{}]*)\}", block)
    return [[int(t, 0) for t in NUM.findall(g)] for g in groups if NUM.findall(g)]


def limbs_to_hex(limbs):
    v = 0
    for i, w in enumerate(limbs):
        v |= w << (64 * i)
    return hex(v)


def array(text, name):
    m = re.search(r"\b" + re.escape(name) + r"\s*(\[[^\]]*\])*\s*=\s*\{", text)
    if not m:
        raise KeyError(name)
    return brace_block(text, m.end() - 1)


def fr_list(block):
    return [limbs_to_hex(g) for g in leaves(block)]


def g1p_list(block):
    """g1p_t initializers are {{x},{y},{z}}: group consecutive coordinate triples."""
    coords = leaves(block)
    assert len(coords) % 3 == 0
    return [[limbs_to_hex(coords[i + j]) for j in range(3)] for i in range(0, len(coords), 3)]


def read(name):
    with open(os.path.join(REF, name)) as f:
        return f.read()


def op_sections(text, prefix):
    """Split a *_kat.cu test body at the '// <op>' markers and parse each section's arrays."""
    body = keep_if_branches(text)
    marks = [(m.start(), m.group(1).strip()) for m in re.finditer(r"^\s*// ((?:%s|Mandelbrot)[^\n]*)$" % prefix, body, re.M)]
    res = {}
    for idx, (pos, title) in enumerate(marks):
        end = marks[idx + 1][0] if idx + 1 < len(marks) else len(body)
        sec = strip_comments(body[pos:end])
        arrays = {}
        for m in re.finditer(r"\b([a-z])\s*\[\s*\]\s*(\[\s*\d+\s*\])?\s*=\s*\{", sec):
            arrays[m.group(1)] = [limbs_to_hex(g) for g in leaves(brace_block(sec, m.end() - 1))]
        key = title.split(",")[0].strip().split()[0].lower()
        res[key] = arrays
    return res


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not found at %s" % REF)

    t = strip_comments(read("fr_fft_testvector.cu"))
    fr_fft = {"q": fr_list(array(t, "q")), "a": fr_list(array(t, "a"))}
    assert len(fr_fft["q"]) == 512 and len(fr_fft["a"]) == 512

    t = strip_comments(read("fk20test_kat.cu"))
    g1_fft = {"input": g1p_list(array(t, "g1p_input")), "output": g1p_list(array(t, "g1p_correct"))}
    assert len(g1_fft["input"]) == 512 and len(g1_fft["output"]) == 512

    t = strip_comments(read("fr_roots.cu"))
    roots = fr_list(array(t, "fr_roots"))
    assert len(roots) == 515

    t = strip_comments(keep_if_branches(read("g1test_kat.cu")))
    g1 = {}
    for name in ["g1p_x0", "g1p_x2", "g1p_x3", "g1p_x24", "g1p_x25", "g1p_wG", "g1p_w8G"]:
        pts = g1p_list(array(t, name))
        assert len(pts) == 1
        g1[name] = pts[0]

    fp_ops = op_sections(read("fptest_kat.cu"), "fp_")
    fr_ops = op_sections(read("frtest_kat.cu"), "fr_")

    fixtures = {
        "fr_fft_kat.json": {"source": "fr_fft_testvector.cu:10-1038", "n": 512, **fr_fft},
        "g1p_fft_kat.json": {"source": "fk20test_kat.cu:31,1317", "n": 512, **g1_fft},
        "fr_roots.json": {"source": "fr_roots.cu:17-533", "roots": roots},
        "g1_kat.json": {"source": "g1test_kat.cu:10-60", **g1},
        "fp_kat.json": {"source": "fptest_kat.cu:18-716", "ops": fp_ops},
        "fr_kat.json": {"source": "frtest_kat.cu:15-650", "ops": fr_ops},
    }
    for fname, data in fixtures.items():
        with open(os.path.join(OUT, fname), "w") as f:
            json.dump(data, f, indent=0, separators=(",", ":"))
            f.write("\n")
        print("wrote", fname)


if __name__ == "__main__":
    main()
