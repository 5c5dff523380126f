"""Writes the library sources (core_amd/csrc) as plain C++ for the CPU
emulation: every KERNEL<<<grid, block, shmem, stream>>>(args); becomes an
emulated launch, and every __shared__ variable registers itself so the
emulator can fill it with garbage before each block (fresh LDS is not
zero)."""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "..", "core_amd", "csrc")
OUT = os.path.join(HERE, "_build", "src")


def split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "(<[":
            depth += 1
        elif ch in ")>]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    parts.append(cur)
    return [p.strip() for p in parts]


def rewrite(text):
    out, i = "", 0
    while True:
        j = text.find("<<<", i)
        if j < 0:
            return out + text[i:]
        k = text.find(">>>", j)
        # kernel name: walk back over identifier and balanced template args
        s = j
        if text[s - 1] == ">":
            depth = 0
            while True:
                s -= 1
                if text[s] == ">":
                    depth += 1
                elif text[s] == "<":
                    depth -= 1
                    if depth == 0:
                        break
        while s > 0 and (text[s - 1].isalnum() or text[s - 1] in "_:"):
            s -= 1
        name = text[s:j]
        cfg = split_top(text[j + 3:k])
        # args: balanced parens after >>>
        a0 = k + 3
        assert text[a0] == "(", text[j - 40:k + 40]
        depth, e = 0, a0
        while True:
            if text[e] == "(":
                depth += 1
            elif text[e] == ")":
                depth -= 1
                if depth == 0:
                    break
            e += 1
        args = text[a0 + 1:e]
        tag = re.sub(r"\s+", " ", name)
        shmem = cfg[2] if len(cfg) > 2 else "0"
        out += text[i:s] + 'emu::launch("%s", dim3(%s), dim3(%s), %s, [=] { %s(%s); })' % (
            tag, cfg[0], cfg[1], shmem, name, args)
        i = e + 1


SHARED = re.compile(r"^(\s*)__shared__\s+([A-Za-z_][\w:]*)\s+([^;]+);", re.M)
DYN_SHARED = re.compile(r"extern\s+__shared__\s+([A-Za-z_][\w:]*)\s+(\w+)\s*\[\s*\]\s*;")
# empty inline-asm register fences (scheduling hints for the GPU compiler)
ASM_FENCE = re.compile(r"^\s*asm volatile\(\"\" : \"\+v\"\([^;]*\)\);.*$", re.M)


def dyn_shared(text):
    text = DYN_SHARED.sub(lambda mo: "%s* const %s = reinterpret_cast<%s*>(emu::dyn_lds());"
                          % (mo.group(1), mo.group(2), mo.group(1)), text)
    return ASM_FENCE.sub("", text)


def reg_shared(text):
    def sub(mo):
        ind, ty, decl = mo.groups()
        names = [d.strip().split("[")[0].strip() for d in decl.split(",")]
        regs = " ".join("emu::lds_reg(&%s, sizeof(%s));" % (n, n) for n in names)
        return "%sstatic %s %s; %s" % (ind, ty, decl, regs)
    return SHARED.sub(sub, text)


def main():
    os.makedirs(OUT, exist_ok=True)
    for f in os.listdir(SRC):
        p = os.path.join(SRC, f)
        if f.endswith(".hip") or f.endswith(".cpp"):
            t = reg_shared(dyn_shared(rewrite(open(p).read())))
            t = t.replace('#include "device.h"', '#include "device.h"')
            open(os.path.join(OUT, f.rsplit(".", 1)[0] + ".cpp"), "w").write(t)
        elif f.endswith(".h"):
            open(os.path.join(OUT, f), "w").write(open(p).read().replace(
                '"../../include/sprgpu.h"', '"sprgpu.h"'))


if __name__ == "__main__":
    main()
