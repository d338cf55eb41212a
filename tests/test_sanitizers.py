"""Sanitizer builds of the host code (CPU only; GPU code is not built with
sanitizers).

* AddressSanitizer + UndefinedBehaviorSanitizer (gcc, -fno-sanitize-recover):
  the glTF importer (with its PNG and JPEG decoders and mip filter) and the scene
  builder (the host half of vgi_scene_create: validation, binned SAH, 4-wide
  collapse, node quantisation, texture tables) run by
  `sanitize/host_fuzz.cpp` on the corrupted-document corpus of the import
  fuzz test (untextured, PNG- and JPEG-textured documents), the reference's own glTF
  assets when present, and seeded triangle soups (clustered, degenerate,
  huge, non-finite, bad indices).
  Every built scene is checked: each triangle in exactly one leaf.
* ThreadSanitizer on the oracle's OpenMP twin (clang with LLVM's OpenMP
  runtime and its Archer tool, which tells ThreadSanitizer about OpenMP's
  synchronisation): `sanitize/oracle_race.c` runs two frames of the G-buffer,
  sample generation (query appends through one atomic counter) and the
  denoiser passes on 4 threads.
"""
from __future__ import annotations

import os
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HERE = Path(__file__).resolve().parent
ROCM = Path("/opt/rocm")
LLVM = ROCM / "llvm"


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.skipif(shutil.which("g++") is None or not (ROCM / "include" / "hip").exists(),
                    reason="needs g++ and the HIP headers")
def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    from gltf_fixture import corrupted_documents, scene_from_mesh
    from restir_harness import make_mesh
    from test_gltf_import import _build

    csrc = ROOT / "violet_amd" / "csrc"
    exe = tmp_path / "host_fuzz"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{ROOT / 'include'}",
           f"-I{csrc}", str(HERE / "sanitize" / "host_fuzz.cpp"), str(csrc / "vgi_gltf.cpp"), str(csrc / "vgi_bvh.cpp"),
           str(csrc / "vgi_image.cpp"),
           str(csrc / "vgi_runtime.cpp"), f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}", "-o", str(exe)]
    rc, out = _run(cmd)
    assert rc == 0, out[-4000:]
    docs = corrupted_documents(_build()[0], tmp_path / "a", seed=99)
    docs += corrupted_documents(scene_from_mesh(make_mesh(300)), tmp_path / "b", seed=7)
    # textured: PNG images (decoding, mips) and texture slots in the damaged files
    docs += corrupted_documents(scene_from_mesh(make_mesh(300, textured=True)), tmp_path / "c", seed=5)
    docs += corrupted_documents(scene_from_mesh(make_mesh(300, textured=True), jpeg=True), tmp_path / "d", seed=6)
    ref = Path("/root/reference/assets")
    docs += [p for p in (ref / "Box" / "Box.gltf", ref / "Duck" / "Duck.gltf") if p.exists()]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    rc, out = _run([str(exe), *map(str, docs)], env=env)
    assert rc == 0 and "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "0 failures" in out and f"{len(docs)} documents" in out, out[-2000:]


def _race_input(path: Path, w=96, h=64):
    from restir_harness import make_mesh, orbit_view
    from violet_amd import synth

    mesh = make_mesh(2000)
    fps = []
    prev = None
    for f in range(2):
        v = orbit_view(f, w, h)
        fps.append(np.frombuffer(bytes(synth.make_frame_params(v, prev)), np.float32))
        prev = v
    mats = np.zeros((len(mesh.albedo), 8), np.float32)
    mats[:, :3] = mesh.albedo
    mats[:, 3] = 1.0
    mats[:, 4] = mesh.metallic
    mats[:, 5] = mesh.roughness
    with open(path, "wb") as fh:
        fh.write(struct.pack("<5I", w, h, len(mesh.positions), len(mesh.indices), len(mats)))
        fh.write(np.concatenate(fps).astype(np.float32).tobytes())
        fh.write(np.ascontiguousarray(mesh.positions, np.float32).tobytes())
        fh.write(np.ascontiguousarray(mesh.indices, np.uint32).tobytes())
        fh.write(np.ascontiguousarray(mesh.tri_material, np.uint32).tobytes())
        fh.write(mats.tobytes())


@pytest.mark.skipif(not (LLVM / "bin" / "clang").exists() or not (LLVM / "lib" / "libarcher.so").exists(),
                    reason="needs ROCm's clang with LLVM OpenMP and Archer")
def test_oracle_openmp_twin_under_thread_sanitizer(tmp_path):
    exe = tmp_path / "oracle_race"
    odir = ROOT / "oracle"
    cmd = [str(LLVM / "bin" / "clang"), "-std=gnu11", "-O1", "-g", "-fopenmp", "-fsanitize=thread",
           "-ffp-contract=off", "-DORACLE_REAL=float", "-DORACLE_F32", f"-I{odir}", str(HERE / "sanitize" / "oracle_race.c"),
           str(odir / "oracle_denoise.c"), str(odir / "oracle_kat.c"), "-lm", f"-Wl,-rpath,{LLVM / 'lib'}", "-o", str(exe)]
    rc, out = _run(cmd)
    assert rc == 0, out[-4000:]
    inp = tmp_path / "in.bin"
    _race_input(inp)
    env = dict(os.environ, OMP_NUM_THREADS="4", OMP_TOOL_LIBRARIES=str(LLVM / "lib" / "libarcher.so"),
               TSAN_OPTIONS="halt_on_error=1 exitcode=66 ignore_noninstrumented_modules=1")
    rc, out = _run([str(exe), str(inp)], env=env)
    assert rc == 0 and "WARNING: ThreadSanitizer" not in out, out[-4000:]
    assert "4 threads" in out, out[-2000:]


PLANTED = {
    "thread": ("#include <stdio.h>\nint main(void) { int c = 0;\n#pragma omp parallel for\n"
               "for (int i = 0; i < 1000; ++i) c += i;\nprintf(\"%d\\n\", c); return 0; }\n", "data race"),
    "address": ("#include <stdlib.h>\nint main(void) { int* p = malloc(4 * sizeof(int)); p[4] = 1;"
                " int r = p[0]; free(p); return r; }\n", "heap-buffer-overflow"),
}


@pytest.mark.parametrize("kind", sorted(PLANTED))
def test_sanitizer_setups_report_a_planted_error(tmp_path, kind):
    """The two set-ups above are not vacuous: each reports a planted error of
    its kind in a few lines of C (an unsynchronised OpenMP sum, a heap write
    one past the end)."""
    src, want = PLANTED[kind]
    c = tmp_path / "planted.c"
    c.write_text(src)
    exe = tmp_path / "planted"
    if kind == "thread":
        if not (LLVM / "lib" / "libarcher.so").exists():
            pytest.skip("needs ROCm's clang with LLVM OpenMP and Archer")
        cmd = [str(LLVM / "bin" / "clang"), "-O1", "-g", "-fopenmp", "-fsanitize=thread", str(c),
               f"-Wl,-rpath,{LLVM / 'lib'}", "-o", str(exe)]
        env = dict(os.environ, OMP_NUM_THREADS="4", OMP_TOOL_LIBRARIES=str(LLVM / "lib" / "libarcher.so"),
                   TSAN_OPTIONS="halt_on_error=1 exitcode=66 ignore_noninstrumented_modules=1")
    else:
        cmd = ["gcc", "-O0", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", str(c), "-o", str(exe)]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    rc, out = _run(cmd)
    assert rc == 0, out[-2000:]
    rc, out = _run([str(exe)], env=env)
    assert rc != 0 and want in out, out[-2000:]
