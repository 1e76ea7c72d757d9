"""Write tests/golden/openmc_lib_bindings.txt: every name the reference's
Python bindings (openmc/lib/*.py) look up in libopenmc.so when openmc.lib is
imported, with the number of argtypes it declares for functions.

The reference's modules are read as text (never imported).  Each line is

    func <name> <n_args>     a function (n_args -1: no argtypes given)
    global <name> <ctype>    a global read through <ctype>.in_dll(_dll, name)

Run:  python tests/golden/make_lib_bindings.py [/path/to/reference]
"""
import glob
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def argtypes_count(text):
    """Number of top-level entries of the list literal `text` ([a, b(c, d)])."""
    depth, n, seen = 0, 0, False
    for ch in text[1:-1]:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        elif ch == "," and depth == 0:
            n += 1
            seen = False
            continue
        if not ch.isspace():
            seen = True
    return n + (1 if seen else 0)


def scan(ref):
    funcs, globs = {}, {}
    for path in sorted(glob.glob(os.path.join(ref, "openmc", "lib", "*.py"))):
        src = open(path).read()
        for name in re.findall(r"_dll\.([A-Za-z_][A-Za-z_0-9]*)", src):
            funcs.setdefault(name, -1)
        for m in re.finditer(r"_dll\.([A-Za-z_][A-Za-z_0-9]*)\.argtypes\s*=\s*(\[)", src):
            start = m.start(2)
            depth = 0
            for k in range(start, len(src)):
                if src[k] == "[":
                    depth += 1
                elif src[k] == "]":
                    depth -= 1
                    if depth == 0:
                        funcs[m.group(1)] = argtypes_count(src[start:k + 1])
                        break
        for ctype, name in re.findall(r"(c_\w+|\(c_char\s*\*\s*256\))\.in_dll\(\s*_dll\s*,\s*['\"](\w+)['\"]", src):
            globs[name] = ctype.replace(" ", "")
        for ctype, name in re.findall(r"_DLLGlobal\(\s*(c_\w+)\s*,\s*['\"](\w+)['\"]", src):
            globs[name] = ctype
    # error codes read by name in openmc/lib/error.py (c_int.in_dll(_dll, s))
    err = open(os.path.join(ref, "openmc", "lib", "error.py")).read()
    for name in re.findall(r"['\"](OPENMC_E_\w+)['\"]", err):
        globs[name] = "c_int"
    for name in ("argtypes", "restype", "errcheck"):
        funcs.pop(name, None)
    return funcs, globs


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    funcs, globs = scan(ref)
    out = os.path.join(HERE, "openmc_lib_bindings.txt")
    with open(out, "w") as f:
        for name in sorted(funcs):
            f.write("func %s %d\n" % (name, funcs[name]))
        for name in sorted(globs):
            f.write("global %s %s\n" % (name, globs[name]))
    print("%d functions, %d globals -> %s" % (len(funcs), len(globs), out))


if __name__ == "__main__":
    main()
