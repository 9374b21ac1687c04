"""Struct layouts computed from the text of C headers (x86-64 System V, LP64).

TEST INFRASTRUCTURE: used only by tests/ (and its golden-file script) to
derive field offsets of the reference's plugin-ABI structs from
libmpm/process.h, which cannot be compiled here (it includes
<suitesparse/cs.h>, absent from this image).  It reads the declarations
as text, without a compiler and without stand-in headers.

Handles what the reference's libmpm headers use: comments, integer
#defines (also as expressions of earlier ones), `typedef struct tag {...}
name;`, members of scalar, pointer, function-pointer, enum, array
(multi-dimensional) and previously defined struct types, several
declarators per line, and `struct timespec` (glibc x86-64: two 8-byte
fields).  Layout rules: each member at the next multiple of its alignment,
the struct padded to a multiple of its largest alignment.
"""
import re

SCALARS = {  # name: (size, align)
    "char": (1, 1), "signed char": (1, 1), "unsigned char": (1, 1),
    "short": (2, 2), "unsigned short": (2, 2),
    "int": (4, 4), "unsigned": (4, 4), "unsigned int": (4, 4),
    "long": (8, 8), "unsigned long": (8, 8), "long long": (8, 8),
    "unsigned long long": (8, 8), "size_t": (8, 8), "ssize_t": (8, 8),
    "float": (4, 4), "double": (8, 8), "long double": (16, 16),
    "struct timespec": (16, 8),
}
POINTER = (8, 8)
ENUM = (4, 4)


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _eval(expr, defines):
    expr = expr.strip()
    for _ in range(10):
        new = re.sub(r"\b[A-Za-z_]\w*\b",
                     lambda m: "(" + defines[m.group(0)] + ")" if m.group(0) in defines
                     else m.group(0), expr)
        if new == expr:
            break
        expr = new
    if not re.fullmatch(r"[\d\s()+\-*/]+", expr):
        raise ValueError(f"not an integer constant: {expr}")
    return int(eval(expr.replace("/", "//"), {"__builtins__": {}}))   # digits, + - * / ( )


class Layout:
    def __init__(self):
        self.defines = {}
        self.structs = {}   # typedef name -> (size, align, [(field, offset)])
        self.enums = set()

    def feed(self, text):
        text = strip_comments(text)
        for m in re.finditer(r"^[ \t]*#[ \t]*define[ \t]+(\w+)[ \t]+([^\n]+)$", text, flags=re.M):
            self.defines[m.group(1)] = m.group(2).strip()
        text = re.sub(r"^[ \t]*#[^\n]*$", " ", text, flags=re.M)
        for m in re.finditer(r"\benum\s+(\w+)\s*\{", text):
            self.enums.add(m.group(1))
        for m in re.finditer(r"typedef\s+struct\s+\w*\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
            self._struct(m.group(2), m.group(1))

    def _type(self, base, ptr):
        if ptr:
            return POINTER
        base = " ".join(base.replace("const", " ").split())
        if base in SCALARS:
            return SCALARS[base]
        if base.startswith("enum ") and base[5:] in self.enums:
            return ENUM
        if base in self.structs:
            size, align, _ = self.structs[base]
            return size, align
        raise ValueError(f"unknown type {base!r}")

    def _struct(self, name, body):
        fields, off, align = [], 0, 1
        for decl in (d.strip() for d in body.split(";")):
            if not decl:
                continue
            fp = re.fullmatch(r"(.+?)\(\s*\*\s*(\w+)\s*\)\s*\(.*\)", decl, flags=re.S)
            if fp:   # function pointer
                items = [(fp.group(2), POINTER, 1)]
            else:
                m = re.fullmatch(r"((?:const\s+)?(?:struct\s+|enum\s+|unsigned\s+|signed\s+|long\s+)*\w+)"
                                 r"\s*(.*)", decl, flags=re.S)
                base, rest = m.group(1), m.group(2)
                items = []
                for d in rest.split(","):
                    d = d.strip()
                    dm = re.fullmatch(r"(\**)\s*(\w+)\s*((?:\[[^\]]+\]\s*)*)", d)
                    if not dm:
                        raise ValueError(f"cannot parse declarator {d!r} in {name}")
                    count = 1
                    for dim in re.findall(r"\[([^\]]+)\]", dm.group(3)):
                        count *= _eval(dim, self.defines)
                    items.append((dm.group(2), self._type(base, len(dm.group(1))), count))
            for fname, (size, al), count in items:
                off = (off + al - 1) // al * al
                fields.append((fname, off))
                off += size * count
                align = max(align, al)
        size = (off + align - 1) // align * align
        self.structs[name] = (size, align, fields)

    def lines(self, name):
        """'name sizeof S' and 'name field offset' lines."""
        size, _, fields = self.structs[name]
        return [f"{name} sizeof {size}"] + [f"{name} {f} {o}" for f, o in fields]


def layout_of(paths):
    lay = Layout()
    for p in paths:
        lay.feed(open(p).read())
    return lay
