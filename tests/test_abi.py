"""CPU tests of the C ABI: the library loads, exports every symbol include/pef.h
declares, its host-side scene builders are bit-identical to the oracle's, and
argument validation rejects bad input before any device work.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pef
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pef.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pef_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", pef.LIB_PATH], check=True, capture_output=True, text=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    L = pef.lib()
    for s in decl:
        assert getattr(L, s) is not None
    assert "gfx950" in pef.version()


def test_host_builders_bit_identical_to_oracle():
    for dim, dx, h in [(3, 0.02, 0.04), (3, 0.005, 0.01), (2, 0.05, 0.12), (3, 0.05, 0.11)]:
        a = pef.mass_calibrated_to_rest(1000.0, dx, h, dim)
        b = O.mass_calibrated_to_rest(1000.0, dx, h, dim)
        assert a == b
        assert pef.mass_from_spacing(1000.0, dx, dim) == O.mass_from_spacing(1000.0, dx, dim)
    for dim, mn, mx, dx, layers in [(3, (0, 0, 0), (1.0, 0.5, 0.5), 0.02, 1), (3, (0.1, 0.05, 0.0), (0.6, 0.8, 0.3), 0.05, 2),
                                    (2, (0.1, 0.05, 0), (0.6, 0.8, 0), 0.05, 2)]:
        P = pef.Particles(dim, 1000.0, 1.0, 2 * dx)
        pef.fill_box(P, (0.013, 0.021, 0.03 if dim == 3 else 0.0), 7, 5, 4, dx)
        pef.jitter_positions(P, 3, 0.01 * dx)
        pef.fill_box_boundary(P, mn, mx, dx, layers)
        Q = O.Particles(dim, 1000.0, 1.0, 2 * dx)
        Q.fill_box((0.013, 0.021, 0.03 if dim == 3 else 0.0), 7, 5, 4, dx)
        Q.jitter(3, 0.01 * dx)
        Q.fill_box_boundary(mn, mx, dx, layers)
        q = Q.arrays()
        assert np.array_equal(P.x.view(np.uint64), q["x"].view(np.uint64))
        assert np.array_equal(P.kind, q["kind"])
        assert np.array_equal(P.psi.view(np.uint64), q["psi"].view(np.uint64))


def test_create_rejects_bad_input_before_device_work():
    L = pef.lib()
    h = C.c_void_p()
    assert L.pef_world_create(None, None, None, C.byref(h)) == pef.ERR_NULL
    parts = pef.Particles(3, 1000.0, 0.001, 0.01)
    pef.fill_box(parts, (0.0, 0.0, 0.0), 2, 2, 2, 0.005)
    params = pef.WcsphParams.water(1000.0, 20.0, (0, 0, -9.81))
    keep = []
    cp = pef._c_particles(parts, keep)
    pp = params.to_c(parts)
    pp.kernel = 1
    assert L.pef_world_create(C.byref(pp), C.byref(cp), None, C.byref(h)) == pef.ERR_UNSUPPORTED
    pp = params.to_c(parts)
    pp.surface_tension_sigma = 0.1
    assert L.pef_world_create(C.byref(pp), C.byref(cp), None, C.byref(h)) == pef.ERR_UNSUPPORTED
    pp = params.to_c(parts)
    pp.h = float("nan")
    assert L.pef_world_create(C.byref(pp), C.byref(cp), None, C.byref(h)) == pef.ERR_ARG
    pp = params.to_c(parts)
    pp.dim = 4
    assert L.pef_world_create(C.byref(pp), C.byref(cp), None, C.byref(h)) == pef.ERR_ARG
    bad = pef.Particles(3, 1000.0, 0.001, 0.01)
    pef.fill_box(bad, (0.0, 0.0, 0.0), 2, 2, 2, 0.005)
    bad.x[3, 1] = np.inf
    cpb = pef._c_particles(bad, keep)
    pp = params.to_c(parts)
    assert L.pef_world_create(C.byref(pp), C.byref(cpb), None, C.byref(h)) == pef.ERR_ARG
    bad.x[3, 1] = 0.0
    bad.kind[2] = 7
    cpb = pef._c_particles(bad, keep)
    assert L.pef_world_create(C.byref(pp), C.byref(cpb), None, C.byref(h)) == pef.ERR_ARG
    empty = pef.Particles(3, 1000.0, 0.001, 0.01)
    cpe = pef._c_particles(empty, keep)
    assert L.pef_world_create(C.byref(pp), C.byref(cpe), None, C.byref(h)) == pef.ERR_ARG
    assert L.pef_world_step(None, 1e-4, 1) == pef.ERR_NULL
    n = C.c_uint64()
    assert L.pef_fill_box(3, None, 1, 1, 1, 0.1, None, 0, C.byref(n)) == pef.ERR_NULL
    o = (C.c_double * 3)(0, 0, 0)
    assert L.pef_fill_box(3, o, 2, 2, 2, 0.1, None, 0, C.byref(n)) == pef.ERR_BOUNDS and n.value == 8


def test_scene_configs_particle_counts():
    from pef import scenes
    def shell(n):
        nx, ny, nz = n
        return 2 * nx * ny + 2 * nx * (nz - 2) + 2 * (ny - 2) * (nz - 2)
    for cfg, nb in [(scenes.CPU_8K, 6252), (scenes.CHECK_64K, 26802), (scenes.GPU_1M, 166722)]:
        dims = tuple(int(round(L / cfg.dx)) + 1 for L in cfg.tank)
        assert shell(dims) == nb
    assert scenes.GPU_1M.n_fluid == 1048576 and scenes.CHECK_64K.n_fluid == 65536


def test_batch_scene_generator_draws_like_the_reference_rng():
    from pef import scenes
    r1, r2 = scenes.Xoshiro256StarStar(1), O.Rng(1)
    for _ in range(2000):
        assert r1.uniform(-3.0, 7.0) == r2.uniform(-3.0, 7.0)
    cfgs = scenes.batch_configs()
    counts = [c.n_fluid for c in cfgs]
    assert len(cfgs) == 256 and sum(counts) == 4327299
    assert (cfgs[0].nx, cfgs[0].ny, cfgs[0].nz) == (31, 27, 27)
    assert min(counts) >= 1800 and max(counts) <= 70000
    for c in cfgs:  # block inside the tank with the 2 dx wall gap
        assert (c.nx - 1 + 4) * c.dx <= c.tank[0] + 1e-12
        assert abs((c.ny - 1 + 4) * c.dx - c.tank[1]) < 1e-9
        assert (c.nz - 1 + 4) * c.dx <= c.tank[2] + 1e-12


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every struct of include/pef.h has a ctypes mirror with the same field
    names, offsets and size, as gcc lays the header out."""
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "pef.h")).read(), flags=re.S)
    structs = {}
    for name, body in re.findall(r"typedef struct (\w+) \{(.*?)\}\s*\1;", txt, flags=re.S):
        fields = []
        for decl in filter(None, (d.strip() for d in body.split(";"))):
            rest = re.match(r"(?:const )?\w+\*?\s+(.*)$", decl).group(1)
            fields += [re.sub(r"\[.*", "", v).strip("* ") for v in rest.split(",")]
        structs[name] = fields
    assert len(structs) >= 7
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pef.h"', "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'  printf("{name} sizeof %zu\\n", sizeof({name}));')
        lines += [f'  printf("{name} {f} %zu\\n", offsetof({name}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    c_layout = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        s, f, v = ln.split()
        c_layout[(s, f)] = int(v)
    for name, fields in structs.items():
        cls = getattr(pef, name)
        assert [f[0] for f in cls._fields_] == fields, name
        assert C.sizeof(cls) == c_layout[(name, "sizeof")], name
        for f in fields:
            assert getattr(cls, f).offset == c_layout[(name, f)], (name, f)
