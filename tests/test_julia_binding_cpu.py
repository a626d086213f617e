"""The Julia module (particleincellcodegolf.jl_amd/julia/PICGolf.jl) cannot run
here (no Julia in the image), so its binding is checked as text: every ccall
names a function a header declares and the library exports, and every struct
mirror lists the C struct's fields in the same order with matching widths."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "particleincellcodegolf.jl_amd", "julia", "PICGolf.jl")
HEADERS = ["pic.h", "pic2d.h", "pic1d2v.h", "pic2d3v.h", "picem1d.h"]

C_TYPES = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64", "pic2d3v_species": "Pic2D3VSpecies"}


def declared():
    out = set()
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        out |= set(re.findall(r"^\s*(?:int|const char\*|void)\s+(pic\w+)\s*\(", txt, re.M))
    return out


def c_struct(name):
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        m = re.search(r"typedef struct " + name + r" \{(.*?)\} " + name + ";", txt, re.S)
        if m:
            fields = []
            for line in m.group(1).splitlines():
                line = re.sub(r"/\*.*?\*/", "", line).strip()
                m2 = re.match(r"(int32_t|int64_t|double|pic2d3v_species)\s+(.*);", line)
                if not m2:
                    continue
                for f in m2.group(2).split(","):
                    f = f.strip()
                    arr = re.match(r"(\w+)\[(\w+)\]", f)
                    if arr and not arr.group(2).isdigit():   # a #define'd length
                        n = re.search(r"#define " + arr.group(2) + r"\s+(\d+)", txt).group(1)
                        arr = re.match(r"(\w+)\[(\d+)\]", f"{arr.group(1)}[{n}]")
                    t = C_TYPES[m2.group(1)]
                    fields.append((arr.group(1), f"NTuple{{{arr.group(2)},{t}}}") if arr else (f, t))
            return fields
    raise KeyError(name)


def jl_struct(name):
    txt = open(JL).read()
    m = re.search(r"^struct " + name + r"\n(.*?)^end", txt, re.S | re.M)
    return [tuple(x.strip() for x in line.split("::")) for line in m.group(1).strip().splitlines()]


def test_every_ccall_is_declared_and_exported():
    import picgolf
    names = set(re.findall(r"ccall\(\(:(\w+), libpicgolf\)", open(JL).read()))
    assert names, "no ccall found"
    assert names <= declared(), names - declared()
    L = picgolf.load()
    for n in names:
        assert hasattr(L, n), n


def test_struct_mirrors_match_the_headers():
    for jl, c in (("PicConfig", "pic_config"), ("Pic2DConfig", "pic2d_config"), ("Pic1D2VConfig", "pic1d2v_config"),
                  ("PicTuning", "pic_tuning"), ("Pic2D3VSpecies", "pic2d3v_species"),
                  ("Pic2D3VConfig", "pic2d3v_config"), ("PicEM1DConfig", "picem1d_config")):
        assert jl_struct(jl) == c_struct(c), jl
