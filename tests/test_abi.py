"""CPU-side checks of the boundary: liboxr.so loads and exports every
symbol include/oxr/pathtrace.h declares; the ctypes mirror has the header's
layout; host-side option rules. No device calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "oxr", h) for h in ("pathtrace.h", "guide.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(oxr_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from oxr import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.EXPORTED) == names


def test_nm_shows_c_linkage():
    from oxr import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for n in declared_functions():
        assert re.search(rf"\bT {n}$", out, re.M), n


def test_struct_layouts_match_header():
    from oxr import _abi as A
    # Compile a tiny program against the header and compare sizeof/offsetof.
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "oxr/pathtrace.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(oxr_material), sizeof(oxr_light),
         sizeof(oxr_scene_desc), sizeof(oxr_camera), sizeof(oxr_pt_options), sizeof(oxr_display),
         sizeof(oxr_ray), sizeof(oxr_hit), sizeof(oxr_pt_stats), sizeof(oxr_bvh_info),
         sizeof(oxr_material_sample));
  printf("%zu %zu %zu\n", offsetof(oxr_material, dispersion), offsetof(oxr_scene_desc, lights),
         offsetof(oxr_pt_stats, launches_anyhit));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "t.c")
        open(c, "w").write(prog)
        exe = os.path.join(td, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(x) for x in out[0].split()]
    mine = [C.sizeof(t) for t in (A.Material, A.Light, A.SceneDesc, A.Camera, A.PtOptions, A.Display, A.Ray,
                                  A.Hit, A.PtStats, A.BvhInfo, A.MaterialSample)]
    assert sizes == mine
    offs = [int(x) for x in out[1].split()]
    assert offs == [A.Material.dispersion.offset, A.SceneDesc.lights.offset, A.PtStats.launches_anyhit.offset]
    assert C.sizeof(A.Ray) == 32 and C.sizeof(A.Hit) == 16


def test_no_device_reports_backend_error():
    import oxr
    if oxr.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(oxr.BackendError):
        oxr.PathTracer(0)


def test_validate_mirrors_reference_ranges():
    import oxr
    o = oxr.RenderOptions(width=4, height=4)
    o.validate()
    for bad in (dict(samples_per_pixel=0), dict(samples_per_pixel=(1 << 20) + 1), dict(max_bounces=1025),
                dict(clamp=-1.0), dict(clamp=float("nan"))):
        b = oxr.RenderOptions(path_trace=oxr.PathTraceOptions(**bad))
        with pytest.raises(oxr.InvalidOptions):
            b.validate()


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "oxideav-render_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", "Makefile")):
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", txt).lower().replace("oracle_", ""), f


def test_guide_struct_layouts_match_header():
    from oxr import _abi as A
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "oxr/guide.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(oxr_guide_config), sizeof(oxr_guide_stats), sizeof(oxr_tune_result),
         offsetof(oxr_guide_config, bounces), offsetof(oxr_guide_stats, infer_ms));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "t.c")
        open(c, "w").write(prog)
        exe = os.path.join(td, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out == [C.sizeof(A.GuideConfig), C.sizeof(A.GuideStats), C.sizeof(A.TuneResult),
                   A.GuideConfig.bounces.offset, A.GuideStats.infer_ms.offset]


def test_texture_struct_layouts_match_header():
    from oxr import _abi as A
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "oxr/pathtrace.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(oxr_scene_desc), sizeof(oxr_texture), sizeof(oxr_texture_binding),
         sizeof(oxr_material_textures), offsetof(oxr_scene_desc, textures), offsetof(oxr_scene_desc, colors),
         offsetof(oxr_material_textures, normal_scale));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "t.c")
        open(c, "w").write(prog)
        exe = os.path.join(td, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out == [C.sizeof(A.SceneDesc), C.sizeof(A.Texture), C.sizeof(A.TextureBinding),
                   C.sizeof(A.MaterialTextures), A.SceneDesc.textures.offset, A.SceneDesc.colors.offset,
                   A.MaterialTextures.normal_scale.offset]
