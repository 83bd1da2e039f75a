"""include/sph.h is a plain C header: a C11 program that includes it compiles
warning-free, pins the reference's struct layouts at compile time, links
against libnprsph.so and calls entry points that fail before any device work
(CPU only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <stddef.h>
#include <string.h>
#include "sph.h"

/* Main.cpp:93-99 (std430 Particle), :110-116 (ConstantsUniform), :118-122 (BoundaryUniform) */
_Static_assert(sizeof(sph_particle_t) == 64, "Particle is 4 vec4");
_Static_assert(offsetof(sph_particle_t, pos) == 0, "pos at offset 0, the renderer's attribute 0");
_Static_assert(offsetof(sph_particle_t, vel) == 16, "vel");
_Static_assert(offsetof(sph_particle_t, force) == 32, "force");
_Static_assert(offsetof(sph_particle_t, extras) == 48, "extras");
_Static_assert(sizeof(sph_constants_t) == 16, "ConstantsUniform");
_Static_assert(sizeof(sph_boundary_t) == 32, "BoundaryUniform");

int main(void)
{
    sph_ctx* ctx = NULL;
    size_t counts[1] = {0};
    if (sph_create_batch(&ctx, 0, counts, NULL, NULL, NULL, NULL) != SPH_ERR_INVALID) return 1;
    if (strstr(sph_create_error(), "scene") == NULL) return 2;
    if (sph_create(NULL, 10, NULL, NULL, NULL, NULL) != SPH_ERR_INVALID) return 3;
    if (sph_size(NULL) != 0 || sph_step(NULL, 1) != SPH_ERR_INVALID) return 4;
    if (strcmp(sph_last_error(NULL), "null context") != 0) return 5;
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_header_compiles_as_c11_and_links(tmp_path, pkg):
    lib = pkg.sph.library_path()
    src = tmp_path / "use_sph.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "use_sph"
    libdir = os.path.dirname(lib)
    cc = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
                         str(src), "-L", libdir, "-lnprsph", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                        capture_output=True, text=True)
    assert cc.returncode == 0, cc.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stderr)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a C++ compiler")
def test_header_links_from_cpp(tmp_path, pkg):
    """The extern "C" guard: a C++ caller (as Main.cpp would be) links unmangled names."""
    lib = pkg.sph.library_path()
    src = tmp_path / "use_sph.cpp"
    src.write_text(PROGRAM.replace("_Static_assert", "static_assert"))
    exe = tmp_path / "use_sph_cpp"
    libdir = os.path.dirname(lib)
    cc = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                         str(src), "-L", libdir, "-lnprsph", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                        capture_output=True, text=True)
    assert cc.returncode == 0, cc.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stderr)
