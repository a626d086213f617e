"""CPU-only checks of the product library: it loads, exports every symbol
include/pic.h declares, and rejects bad configurations before touching a GPU."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_functions(name="pic.h", prefix="pic_"):
    txt = open(os.path.join(ROOT, "include", name)).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void)\s+(" + prefix + r"\w+)\s*\(", txt, re.M)))


def test_header_and_binding_agree():
    import picgolf
    assert sorted(picgolf.EXPORTS) == header_functions()


def test_library_exports_every_symbol():
    import picgolf
    L = picgolf.load()
    for name in header_functions():
        assert hasattr(L, name), name
    assert L.pic_version() >= 1
    assert L.pic_strerror(-4) == b"non-finite charge or field"


@pytest.mark.parametrize("field,value,msg", [
    ("N", 100, "power of two"),
    ("N", 8192, "power of two"),
    ("P", 0, "P must"),
    ("shape", 3, "shape"),
    ("scheme", 5, "scheme"),
    ("precision", 7, "precision"),
    ("w", -1.0, "w must"),
    ("max_sweeps", 0, "max_sweeps"),
])
def test_create_rejects_bad_config_with_message(field, value, msg):
    import picgolf
    from picgolf.core import Config
    cfg = Config(N=64, P=2048, W=100.0, dt=1 / 384)
    setattr(cfg, field, value)
    with pytest.raises(picgolf.PicError) as e:
        picgolf.Context(cfg)
    assert e.value.code == -1 and msg in str(e.value)


@pytest.mark.parametrize("kw,msg", [
    (dict(scheme="simpson", shape="ngp"), "Simpson scheme needs"),
    (dict(scheme="simpson", N=2048), "N <= 1024"),
    (dict(scheme="leapfrog", shape="ngp", precision="mixed"), "mixed precision"),
    (dict(scheme="midpoint", shape="area"), "Gaussian"),
    (dict(scheme="leapfrog", shape="area"), "Simpson"),
])
def test_simpson_and_area_combinations_are_checked(kw, msg):
    import picgolf
    from picgolf.core import Config
    base = dict(N=64, P=64, W=100.0, dt=1e-3, shape="gauss7", scheme="simpson")
    base.update(kw)
    with pytest.raises(picgolf.PicError) as e:
        picgolf.Context(Config(**base))
    assert e.value.code == -1 and msg in str(e.value)


def test_midpoint_needs_gaussian_shape():
    import picgolf
    from picgolf.core import Config
    with pytest.raises(picgolf.PicError, match="Gaussian"):
        picgolf.Context(Config(N=64, P=64, W=100.0, dt=1e-3, shape="ngp", scheme="midpoint"))


def test_quiet_start_matches_oracle(orc):
    from picgolf.configs import quiet_start
    x, v = quiet_start(2048)
    xo, vo = orc.quiet_start(2048)
    assert np.array_equal(x, xo) and np.array_equal(v, vo)


def test_presets_follow_the_scripts():
    from picgolf import configs
    cfg, x, v, T = configs.ngp_two_stream()
    assert (cfg.N, cfg.P, cfg.dt, cfg.W, cfg.w, T) == (128, 8192, 1 / 512, 200.0, 3.125, 1024)
    cfg, x, v, T = configs.quiet_two_stream()
    assert (cfg.N, cfg.P, cfg.dt, T) == (64, 2048, 1 / 384, 8192)
    assert cfg.rtol == 4 * np.finfo(float).eps
    cfg, x, v, T = configs.scaled_two_stream(P=4096)
    assert (cfg.N, cfg.dt, cfg.W, T) == (4096, 1 / 24576, 400.0, 2000)
    assert abs(np.std(v[:2048] + 1) - 0.05) < 0.005


def test_header2d_and_binding_agree():
    import picgolf
    assert sorted(picgolf.EXPORTS_2D) == header_functions("pic2d.h", "pic2d_")


def test_library_exports_every_2d_symbol():
    import picgolf
    L = picgolf.load()
    for name in header_functions("pic2d.h", "pic2d_"):
        assert hasattr(L, name), name


@pytest.mark.parametrize("field,value,msg", [
    ("NX", 100, "powers of two"),
    ("NY", 2048, "powers of two"),
    ("P", 0, "P must"),
    ("P", 2 ** 22 + 1, "P must"),
    ("precision", 7, "precision"),
    ("pairing", 3, "pairing"),
    ("dt", 0.0, "dt and w"),
    ("w", float("nan"), "dt and w"),
    ("ns", 0, "ns must"),
])
def test_create2d_rejects_bad_config_with_message(field, value, msg):
    import picgolf
    from picgolf.core2d import Config2D
    cfg = Config2D(NX=32, NY=16, P=4096, dt=1 / 256, w=1.0, B0=1.0)
    setattr(cfg, field, value)
    with pytest.raises(picgolf.PicError) as e:
        picgolf.Context2D(cfg)
    assert e.value.code == -1 and msg in str(e.value)


def test_es2d3v_preset_follows_the_script():
    from picgolf import configs
    cfg, (x, y, vx, vy, vz), T = configs.es2d3v(NX=32, NY=32, ppc=16)
    n0 = 4 * np.pi ** 2
    assert (cfg.NX, cfg.NY, cfg.P, cfg.dt, cfg.ns, T) == (32, 32, 32 * 32 * 16, 1 / 1024, 2, 500)
    assert cfg.B0 == np.sqrt(n0) / 4 and cfg.w == n0 / cfg.P * 32 * 32
    assert np.all((x > 0) & (x <= 1)) and np.all((y > 0) & (y <= 1))
    assert abs(np.std(vx[: cfg.P // 2] + 1) - 0.1) < 0.01 and abs(np.std(vz) - 0.1) < 0.01


def test_header1d2v_and_binding_agree():
    import picgolf
    assert sorted(picgolf.EXPORTS_1D2V) == header_functions("pic1d2v.h", "pic1d2v_")


def test_library_exports_every_1d2v_symbol():
    import picgolf
    L = picgolf.load()
    for name in header_functions("pic1d2v.h", "pic1d2v_"):
        assert hasattr(L, name), name


@pytest.mark.parametrize("field,value,msg", [
    ("N", 100, "power of two"),
    ("shape", 1, "shape"),
    ("bin", 0, "bin"),
    ("P", 0, "P must"),
    ("n0", 0.0, "n0"),
    ("q", (1.0, 2.0, 3.0), "nspecies"),
])
def test_create1d2v_rejects_bad_config_with_message(field, value, msg):
    import ctypes
    import picgolf
    from picgolf.core1d2v import Config1D2V
    cfg = Config1D2V(N=64, P=640, bin=4, w=0.1, n0=39.0, B0=1.0, dt=0.01)
    if field == "q":
        c = cfg.to_c()
        c.nspecies = 3
        h = ctypes.c_void_p()
        L = picgolf.load()
        assert L.pic1d2v_create(ctypes.byref(h), ctypes.byref(c)) == -1
        assert msg in L.pic1d2v_last_error(None).decode()
        return
    setattr(cfg, field, value)
    with pytest.raises(picgolf.PicError) as e:
        picgolf.Context1D2V(cfg)
    assert e.value.code == -1 and msg in str(e.value)


def test_header2d3v_and_binding_agree():
    import picgolf
    assert sorted(picgolf.EXPORTS_2D3V) == header_functions("pic2d3v.h", "pic2d3v_")


def test_library_exports_every_2d3v_symbol():
    import picgolf
    L = picgolf.load()
    for name in header_functions("pic2d3v.h", "pic2d3v_"):
        assert hasattr(L, name), name


def test_2d3v_struct_layout_matches_the_header(tmp_path):
    """The ctypes mirrors have the C structs' sizes and field offsets."""
    import ctypes
    import subprocess
    from picgolf import _lib as L
    structs = {"pic2d3v_config": L.pic2d3v_config, "pic2d3v_species": L.pic2d3v_species,
               "pic1d2v_config": L.pic1d2v_config}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pic2d3v.h"', '#include "pic1d2v.h"',
             "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for name, cls in structs.items():
        assert got[(name, "size")] == ctypes.sizeof(cls), name
        for f, _ in cls._fields_:
            assert got[(name, f)] == getattr(cls, f).offset, (name, f)


@pytest.mark.parametrize("change,msg", [
    (dict(NX=100), "powers of two"),
    (dict(Lx=0.0), "Lx and Ly"),
    (dict(dt=-1.0), "dt must"),
    (dict(B0=(0.0, float("nan"), 0.0)), "B0 must"),
    (dict(ntskip=0), "ntskip"),
    (dict(ngskip=3), "ngskip"),
    (dict(field_update=5), "field_update"),
    (dict(boris=9), "boris"),
    (dict(species=[]), "nspecies"),
    (dict(species=[dict(P=0)]), "P must"),
    (dict(species=[dict(shape=9)]), "shape"),
    (dict(species=[dict(mass=0.0)]), "mass"),
    (dict(species=[dict(P=2 ** 25), dict(P=2 ** 25), dict(P=1)]), "2^26"),
    (dict(field="lorenz_semi_implicit", maxiters=-1), "maxiters"),
    (dict(field="lorenz_semi_implicit", rtol=-1e-8), "rtol"),
    (dict(field="lorenz_semi_implicit", theta_source=float("inf")), "theta_source"),
    # the buffered deposit grid must hold every stencil cell (the reference's BoundsError)
    (dict(field="lorenz_semi_implicit", buffer=1, species=[dict(shape=4)]), "buffer >= 2"),
    (dict(field="lorenz_semi_implicit", buffer=2, species=[dict(shape=1), dict(shape=7)]), "buffer >= 3"),
    (dict(precision=7), "precision"),
    (dict(precision="mixed", field="lorenz_staggered"), "mixed precision"),
])
def test_create2d3v_rejects_bad_config_with_message(change, msg):
    import picgolf
    from picgolf.core2d3v import Config2D3V, SpeciesConfig
    kw = dict(NX=16, NY=8, dt=0.01, species=[SpeciesConfig(P=100)])
    for k, v in change.items():
        kw[k] = [SpeciesConfig(**{"P": 100, **s}) for s in v] if k == "species" else v
    with pytest.raises(picgolf.PicError) as e:
        picgolf.Context2D3V(Config2D3V(**kw))
    assert e.value.code == -1 and msg in str(e.value)


def test_header_em1d_and_binding_agree():
    import picgolf
    assert sorted(picgolf.EXPORTS_EM1D) == header_functions("picem1d.h", "picem1d_")


def test_library_exports_every_em1d_symbol():
    import picgolf
    L = picgolf.load()
    for name in header_functions("picem1d.h", "picem1d_"):
        assert hasattr(L, name), name


def test_em1d_struct_layout_matches_the_header(tmp_path):
    import ctypes
    import subprocess
    from picgolf import _lib as L
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "picem1d.h"', "int main(void) {",
             'printf("size %zu\\n", sizeof(picem1d_config));']
    for f, _ in L.picem1d_config._fields_:
        lines.append(f'printf("{f} %zu\\n", offsetof(picem1d_config, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.split("\n") if l)
    assert int(got["size"]) == ctypes.sizeof(L.picem1d_config)
    for f, _ in L.picem1d_config._fields_:
        assert int(got[f]) == getattr(L.picem1d_config, f).offset, f


@pytest.mark.parametrize("change,msg", [
    (dict(N=100), "power of two"),
    (dict(N=2048), "power of two"),
    (dict(P=0), "P must"),
    (dict(max_sweeps=0), "max_sweeps"),
    (dict(eps=0.0), "eps and dt"),
    (dict(dt=float("inf")), "eps and dt"),
    (dict(rtol=-1.0), "rtol"),
    (dict(scheme="explicit", N=48), "power of two"),
    (dict(scheme="explicit", dt=0.0), "eps and dt"),
])
def test_create_em1d_rejects_bad_config_with_message(change, msg):
    import picgolf
    from picgolf.core_em1d import ConfigEM1D
    kw = dict(N=64, P=1024, eps=0.1, dt=1e-4, w=25.0)
    kw.update(change)
    with pytest.raises(picgolf.PicError) as e:
        picgolf.ContextEM1D(ConfigEM1D(**kw))
    assert e.value.code == -1 and msg in str(e.value)


def test_create_em1d_rejects_unknown_scheme():
    import picgolf
    from picgolf import _lib as L
    from picgolf.core_em1d import ConfigEM1D
    with pytest.raises(ValueError):
        ConfigEM1D(N=64, P=1024, eps=0.1, dt=1e-4, w=25.0, scheme="implicit").to_c()
    c = ConfigEM1D(N=64, P=1024, eps=0.1, dt=1e-4, w=25.0).to_c()
    c.scheme = 7
    lib = L.load()
    import ctypes
    h = ctypes.c_void_p()
    assert lib.picem1d_create(ctypes.byref(h), ctypes.byref(c)) == -1
    assert b"scheme" in lib.picem1d_last_error(None)
