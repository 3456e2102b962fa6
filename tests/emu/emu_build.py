"""Host (CPU) build of a HIP source of this package against the emulator in
tests/emu/hip/hip_runtime.h, for AddressSanitizer runs of the kernels (test infrastructure).

The source is copied with two mechanical rewrites that g++ needs and that change no kernel
logic:
* ``kernel<targs><<<grid, block, shmem, stream>>>(args);`` becomes
  ``emu::launch(grid, block, [&]() { kernel<targs>(args); });``
* the empty ``asm volatile("" : "+v"(x))`` optimisation barriers get a host register class.

    python tests/emu/emu_build.py SOURCE DRIVER OUT [--no-asan]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _match_back(s, i, open_c, close_c):
    """s[i] == close_c; returns the index of the matching open_c."""
    depth = 0
    while i >= 0:
        if s[i] == close_c:
            depth += 1
        elif s[i] == open_c:
            depth -= 1
            if depth == 0:
                return i
        i -= 1
    raise ValueError("unbalanced %s%s" % (open_c, close_c))


def _match_fwd(s, i, open_c, close_c):
    depth = 0
    while i < len(s):
        if s[i] == open_c:
            depth += 1
        elif s[i] == close_c:
            depth -= 1
            if depth == 0:
                return i
        i += 1
    raise ValueError("unbalanced %s%s" % (open_c, close_c))


def _split_top(s):
    out, depth, cur = [], 0, ""
    for c in s:
        if c in "(<[{":
            depth += 1
        elif c in ")>]}":
            depth -= 1
        if c == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += c
    if cur.strip():
        out.append(cur.strip())
    return out


def rewrite_launches(src):
    out = []
    pos = 0
    while True:
        k = src.find("<<<", pos)
        if k < 0:
            out.append(src[pos:])
            break
        # callee: identifier with an optional template argument list just before <<<
        j = k - 1
        if src[j] == ">":
            j = _match_back(src, j, "<", ">") - 1
        m = re.search(r"[A-Za-z_][A-Za-z_0-9:]*$", src[:j + 1])
        start = m.start()
        callee = src[start:k]
        e = src.find(">>>", k)
        cfg = _split_top(src[k + 3:e])
        p0 = src.index("(", e)
        p1 = _match_fwd(src, p0, "(", ")")
        args = src[p0 + 1:p1]
        out.append(src[pos:start])
        out.append("emu::launch(dim3(%s), dim3(%s), [&]() { %s(%s); })"
                   % (cfg[0], cfg[1], callee, args))
        pos = p1 + 1
    return "".join(out)


def transform(text):
    text = rewrite_launches(text)
    text = text.replace('"+v"', '"+r"')
    return text


def build(source, driver, out, asan=True, extra=()):
    csrc = os.path.dirname(os.path.abspath(source))
    bdir = os.path.dirname(os.path.abspath(out))
    os.makedirs(bdir, exist_ok=True)
    tsrc = os.path.join(bdir, os.path.basename(source) + ".emu.cpp")
    with open(source) as f:
        text = transform(f.read())
    with open(tsrc, "w") as f:
        f.write('#line 1 "%s"\n' % os.path.abspath(source))
        f.write(text)
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
             "-I" + HERE, "-I" + csrc, "-Wno-unused-result"]
    if asan:
        flags += ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=all"]
    cmd = ["g++"] + flags + list(extra) + [tsrc, os.path.abspath(driver), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("emulator build failed:\n" + r.stderr[-6000:])
    return out


if __name__ == "__main__":
    a = [x for x in sys.argv[1:] if not x.startswith("--")]
    print(build(a[0], a[1], a[2], asan="--no-asan" not in sys.argv))
