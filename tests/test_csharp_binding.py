"""The C# drop-in (ddla_amd/csharp/*.cs) against the C ABI: every DllImport binds an entry
point that include/ddla_native.h declares, with the same number of arguments and a C# type
that marshals each C argument (64-bit dimensions as long, int flags as int or the reference's
flag enums, pointers by IntPtr / ref / in / out / arrays); HipNative.cs declares every entry
point of the header; the retargeted Source partial keeps the BLIS argument lists of
DDLA/BLAS/Native/3m.cs.  No .NET toolchain is needed: the files are parsed as text."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CS_DIR = os.path.join(ROOT, "ddla_amd", "csharp")
HEADER = os.path.join(ROOT, "include", "ddla_native.h")


def header_prototypes():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"^(int|const char\*|void)\s+(ddla_\w+)\((.*?)\);", src,
                                      flags=re.S | re.M):
        args = " ".join(args.split())
        params = []
        if args and args != "void":
            for a in args.split(","):
                m = re.match(r"(.*?)([A-Za-z_]\w*)$", a.strip())
                params.append((" ".join(m.group(1).replace("*", " * ").split()), m.group(2)))
        out[name] = params
    return out


def cs_imports():
    decls = []
    for fn in sorted(os.listdir(CS_DIR)):
        if not fn.endswith(".cs"):
            continue
        text = open(os.path.join(CS_DIR, fn)).read()
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r'\[DllImport\([^\]]*EntryPoint\s*=\s*"(\w+)"\)\]\s*'
                             r'(?:public|private|internal)\s+static\s+extern\s+(\w+)\s+(\w+)\s*'
                             r'\((.*?)\);', text, flags=re.S):
            entry, ret, method, params = m.groups()
            params = " ".join(params.split())
            plist = [p.strip() for p in params.split(",")] if params else []
            decls.append((fn, entry, ret, method, plist))
    return decls


VALUE = {
    "ddla_dim_t": {"long", "dim_t"},
    "ddla_inc_t": {"long", "inc_t"},
    "int64_t": {"long"},
    "int": {"int", "TransType", "UpLo", "SideType", "DiagType", "ConjType"},
    "int32_t": {"int"},
    "uint64_t": {"ulong"},
    "double": {"double"},
    "size_t": {"nuint"},
    "ddla_batch_t": {"IntPtr"},
    "hipStream_t": {"IntPtr"},
}
POINTER_FORMS = ("IntPtr", "ref ", "in ", "out ", "long[]", "string", "StringBuilder")


def test_every_import_matches_the_header():
    protos = header_prototypes()
    decls = cs_imports()
    assert len(decls) > 60
    for fn, entry, ret, method, plist in decls:
        assert entry in protos, f"{fn}: {method} binds {entry}, not declared in the header"
        cparams = protos[entry]
        assert len(plist) == len(cparams), \
            f"{fn}: {method} ({entry}) takes {len(plist)} arguments, the header {len(cparams)}"
        assert ret in ("int", "IntPtr", "void"), f"{fn}: {method} returns {ret}"
        for cs, (ctype, cname) in zip(plist, cparams):
            cs_type = cs.rsplit(" ", 1)[0]
            if "*" in ctype:
                assert cs_type.startswith(POINTER_FORMS), \
                    f"{fn}: {method} argument {cname} ({ctype}) declared as {cs_type}"
            else:
                allowed = VALUE.get(ctype)
                assert allowed is not None, f"unmapped C type {ctype}"
                assert cs_type in allowed, \
                    f"{fn}: {method} argument {cname} ({ctype}) declared as {cs_type}"


def test_hipnative_covers_the_header():
    protos = header_prototypes()
    bound = {entry for fn, entry, *_ in cs_imports() if fn == "HipNative.cs"}
    missing = sorted(set(protos) - bound)
    assert not missing, f"HipNative.cs lacks {missing}"


# argument counts of the BLIS level-3 declarations of DDLA/BLAS/Native/3m.cs that the Source
# partial retargets (method name -> count)
BLIS_ARITY = {"GeMM": 16, "GeMMT": 16, "SyMM": 17, "SyRk": 12, "SyR2k": 16, "TrMM": 13,
              "TrMM3": 18, "TrSM": 13}


@pytest.mark.parametrize("method,arity", sorted(BLIS_ARITY.items()))
def test_source_partial_keeps_the_blis_argument_lists(method, arity):
    decls = {m: (e, p) for fn, e, r, m, p in cs_imports() if fn == "Source.Hip.cs"}
    assert method in decls
    entry, plist = decls[method]
    assert len(plist) == arity
    assert plist[0].split()[0] in ("TransType", "UpLo", "SideType")
    assert any(p.startswith("in ctype alpha") for p in plist)
